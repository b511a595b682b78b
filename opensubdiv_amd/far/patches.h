// Feature-adaptive patch tables for Catmark meshes (manifold, quad or n-gon
// base faces, semi-sharp creases, boundaries with regular boundary points),
// their patch-point stencils, the patch
// map and limit stencil tables -- the inputs of EvalPatches and of
// derivative EvalStencils.
//
// Patches (far/patchTableFactory.cpp:915-1320, far/patchBuilder.cpp:581-700):
// a face of level d < isoLevel is refined when it touches an extraordinary
// vertex or a vertex with a sharp vertex or incident sharp edge (the
// feature test of far/topologyRefiner.cpp:629-700 for this scope); the
// children of refined faces that are not refined again become patches at
// their depth.  Patches are bicubic B-spline unless a corner is
// extraordinary, which can only happen at isoLevel, where the face becomes a
// Gregory-basis end cap (far/catmarkPatchBuilder.cpp:602-1416).
//
// Point numbering of the vertex buffer the tables index:
//   [ control vertices | refined vertices used by patches | end-cap points ]
// The patch-point stencil table holds one row per refined vertex and per
// end-cap point, factorized to control vertices (what Far produces with
// AppendLocalPointStencilTable, far/stencilTableFactory.cpp:256-408).
#pragma once
#include <cstdint>
#include <vector>

#include "stencils.h"

namespace osdfar {

struct PatchArrayRec {
    int regDesc, desc, numPatches, indexBase, stride, primitiveIdBase;
};
struct PatchParamRec {
    uint32_t field0, field1;
    float sharpness;
};
struct PatchHandleRec {
    int arrayIndex, patchIndex, vertIndex;
};

// Options of Far::PatchTableFactory / RefineAdaptive the compiler follows.
struct PatchOptions {
    // smooth extraordinary vertices are isolated to this level only (sharp
    // features to isoLevel): AdaptiveOptions::secondaryLevel
    int secondaryLevel = 15;
    // smooth single-face boundary corners patched as sharp corners
    // (generateLegacySharpCornerPatches, Far's default) or given end caps
    bool legacySharpCorners = true;
    // end-cap points shared between neighbouring caps (shareEndCapPatchPoints)
    bool shareEndCapPoints = true;
};

class PatchTable {
public:
    // Returns false for meshes outside the compiler's scope (non-manifold
    // vertices, n-gon base faces with isoLevel 0).  `levels`, when given,
    // holds the vertex-mode level stencils of the refiner up to at least
    // isoLevel (computed here otherwise).
    bool build(Refiner &refiner, int isoLevel,
               const std::vector<LevelStencils> *levels = nullptr,
               const PatchOptions &opt = PatchOptions());

    int numControl = 0, numRefined = 0, numLocal = 0, maxDepth = 0;
    std::vector<PatchArrayRec> arrays;      // [REGULAR (16), GREGORY_BASIS (20)]
    std::vector<int> indices;               // into the vertex buffer above
    std::vector<PatchParamRec> params;
    LevelStencils pointStencils;            // rows: refined vertices then end-cap points

    // PatchMap (far/patchMap.h:58-231): per ptex face a quadtree; a child is
    // a node (>= 0) or a patch (-(index + 1)); -INT_MAX marks an empty child.
    struct Node { int child[4]; };
    std::vector<int> ptexRoot;
    std::vector<Node> nodes;
    bool findPatch(int ptex, double u, double v, PatchHandleRec &h) const;

    // Far::PatchTable::EvaluateBasis-style weights of patch `p` at ptex
    // (u, v): returns the point count, fills wP/wDu/wDv (and 2nd derivatives
    // when non-null), derivatives scaled to ptex space.
    int evalBasis(int p, double u, double v, double *wP, double *wDu, double *wDv,
                  double *wDuu = nullptr, double *wDuv = nullptr, double *wDvv = nullptr) const;
    const int *patchPoints(int p, int &n) const;

    // Far::StencilTable of the patch points (control-vertex stencils are
    // not included: control vertices sit at the start of the buffer).
    template <typename REAL>
    void pointStencilTable(StencilTableT<REAL> &out) const;

    // LimitStencilTableFactory equivalent: one stencil (P, du, dv[, duu,
    // duv, dvv]) per (ptex, u, v) location, factorized to control vertices.
    // Locations outside every patch get no stencil (as in Far); returns their count.
    template <typename REAL>
    int limitStencils(int n, const int *ptex, const double *u, const double *v, bool second,
                      StencilTableT<REAL> &out) const;

    // Face-varying channels (fvar.cpp): linear refinement and linear QUADS
    // patches, Far's defaults.  The channel's buffer is
    // [base values | refined values], its stencils give the refined values
    // from the base ones; arrays / params mirror the vertex patches.
    struct FVarChannel {
        int numBase = 0, numRefined = 0;
        std::vector<PatchArrayRec> arrays;
        std::vector<int> indices;
        std::vector<PatchParamRec> params;
        LevelStencils stencils;
    };
    std::vector<FVarChannel> fvar;
    // Adds a channel: numValues values, one value index per face-vertex of
    // the base level.  False for out-of-range indices.
    bool buildFaceVarying(const Refiner &r, int numValues, const int *valueIndices);
    template <typename REAL>
    void fvarStencilTable(int channel, StencilTableT<REAL> &out) const;

private:
    std::vector<int> patchDepth_;           // per patch (for derivative scaling)
    std::vector<std::pair<int, int>> patchLeaf_;   // per patch: (level, face)
};

// One (ptex, u, v) per vertex of a uniform refinement level: the location
// of the vertex on the limit surface parameterization (corner of a refined
// face).  Level must be refined already.
void uniformVertexLocations(const Refiner &r, int level, std::vector<int> &ptex,
                            std::vector<double> &u, std::vector<double> &v);

}  // namespace osdfar
