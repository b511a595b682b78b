// Factorized stencil tables (every refined vertex as a weighted sum of
// control vertices), the output format of Far::StencilTableFactoryReal
// (far/stencilTable.h:141-266): sizes / offsets / indices / weights, CSR.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "topology.h"

namespace osdfar {

// Allocator that leaves new elements uninitialised, so large row arrays are
// first touched by the threads that fill them.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U> struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U> &) {}
    template <class U> void construct(U *p) noexcept { ::new ((void *)p) U; }
    template <class U, class... A> void construct(U *p, A &&...a)
    {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};

template <class T>
using hvec = std::vector<T, NoInitAlloc<T>>;

template <typename REAL>
struct StencilTableT {
    int numControlVertices = 0;
    hvec<int> sizes;
    hvec<int64_t> offsets;     // 64-bit: batched tables exceed 2^31 entries
    hvec<int> indices;
    hvec<REAL> weights;
    // optional derivative weights (limit stencils), same shape as weights
    hvec<REAL> du, dv, duu, duv, dvv;
    int numStencils() const { return (int)sizes.size(); }
    size_t nnz() const { return indices.size(); }
};

using StencilTable = StencilTableT<float>;
using StencilTableD = StencilTableT<double>;

enum InterpolationMode { INTERPOLATE_VERTEX = 0, INTERPOLATE_VARYING = 1 };

struct StencilOptions {
    int interpolationMode = INTERPOLATE_VERTEX;
    bool generateIntermediateLevels = false;
    bool generateControlVerts = false;
};

// One sparse row per vertex of a level, entries sorted by control index.
struct LevelStencils {
    std::vector<int64_t> off;   // nv + 1
    hvec<int> idx;
    hvec<double> w;
    int size(int v) const { return (int)(off[v + 1] - off[v]); }
};

// Host threads used by the table compiler (0 = all available).  Callers
// that build many meshes concurrently set 1 per worker.
void setCompilerThreads(int n);
int compilerThreads();

// Computes factorized stencils of every level 1..maxLevel of the refiner.
// levelStencils[l] holds level l (levelStencils[0] is empty: identity).
void computeLevelStencils(const Refiner &r, int mode, std::vector<LevelStencils> &levelStencils);

// Far::StencilTableFactory::Create equivalent (stencils sorted by control
// index inside each row, weights accumulated in double then rounded).
template <typename REAL>
void buildStencilTable(const Refiner &r, const StencilOptions &opt, StencilTableT<REAL> &out);

// The same table from level stencils already computed for the refiner
// (computeLevelStencils, vertex mode; levels.size() > r.maxLevel()).
template <typename REAL>
void stencilTableFromLevels(const Refiner &r, const std::vector<LevelStencils> &levels,
                            const StencilOptions &opt, StencilTableT<REAL> &out);

}  // namespace osdfar
