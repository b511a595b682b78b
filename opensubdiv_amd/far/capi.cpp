// extern "C" layer of libosdfar (include/osdfar.h).
#include <cstring>
#include <memory>
#include <new>

#include "../../include/osdfar.h"
#include "patches.h"

using namespace osdfar;

struct osdfar_topology {
    Refiner refiner;
    std::vector<LevelStencils> levels;     // cached vertex-mode level stencils
    int fvarValues = 0;                    // face-varying channel (0: none)
    std::vector<int> fvarIndices;          // one value index per base face-vertex
};

struct osdfar_stencil_table {
    bool isDouble = false;
    StencilTable f;
    StencilTableD d;
    int numWeightArrays = 1;
};

struct osdfar_patch_table {
    PatchTable table;
};

namespace {
template <typename REAL>
static int copyTable(const StencilTableT<REAL> &s, int *sizes, int *offsets32, long long *offsets64,
                     int *indices, int which, float *w, double *wd)
{
    const hvec<REAL> *arr[6] = { &s.weights, &s.du, &s.dv, &s.duu, &s.duv, &s.dvv };
    if (which < 0 || which > 5) return -1;
    if (offsets32 && s.nnz() >= (size_t)0x7fffffff) return -2;
    const size_t n = s.sizes.size();
    if (sizes) std::memcpy(sizes, s.sizes.data(), n * sizeof(int));
    if (offsets32) for (size_t i = 0; i < n; ++i) offsets32[i] = (int)s.offsets[i];
    if (offsets64) for (size_t i = 0; i < n; ++i) offsets64[i] = (long long)s.offsets[i];
    if (indices) std::memcpy(indices, s.indices.data(), s.nnz() * sizeof(int));
    const hvec<REAL> &src = *arr[which];
    if ((w || wd) && src.size() != s.nnz()) return -3;
    if (w) for (size_t i = 0; i < src.size(); ++i) w[i] = (float)src[i];
    if (wd) for (size_t i = 0; i < src.size(); ++i) wd[i] = (double)src[i];
    return 0;
}

}  // namespace

extern "C" {

int osdfar_set_threads(int n)
{
    setCompilerThreads(n);
    return 0;
}

osdfar_topology *osdfar_topology_create(int numVertices, int numFaces,
                                        const int *numVertsPerFace,
                                        const int *vertIndicesPerFace,
                                        int numCreases, const int *creaseVertexIndexPairs,
                                        const float *creaseWeights,
                                        int numCorners, const int *cornerVertexIndices,
                                        const float *cornerWeights,
                                        int boundaryInterpolation)
{
    try {
        if (numVertices <= 0 || numFaces <= 0 || !numVertsPerFace || !vertIndicesPerFace) return nullptr;
        TopologyDesc d;
        d.numVertices = numVertices;
        d.numFaces = numFaces;
        d.numVertsPerFace = numVertsPerFace;
        d.vertIndicesPerFace = vertIndicesPerFace;
        d.numCreases = creaseVertexIndexPairs && creaseWeights ? numCreases : 0;
        d.creaseVertexIndexPairs = creaseVertexIndexPairs;
        d.creaseWeights = creaseWeights;
        d.numCorners = cornerVertexIndices && cornerWeights ? numCorners : 0;
        d.cornerVertexIndices = cornerVertexIndices;
        d.cornerWeights = cornerWeights;
        d.boundaryInterpolation = boundaryInterpolation;
        std::unique_ptr<osdfar_topology> t(new (std::nothrow) osdfar_topology);
        if (!t || !t->refiner.create(d)) return nullptr;
        return t.release();
    } catch (...) {
        return nullptr;
    }
}

int osdfar_topology_set_creasing(osdfar_topology *topo, int method)
{
    if (!topo || (method != CREASE_UNIFORM && method != CREASE_CHAIKIN)) return -1;
    if (topo->refiner.maxLevel() > 0 || !topo->levels.empty()) return -2;
    topo->refiner.creasing = method;
    return 0;
}

int osdfar_topology_set_scheme(osdfar_topology *topo, int scheme)
{
    if (!topo) return -1;
    if (topo->refiner.maxLevel() > 0 || !topo->levels.empty()) return -2;
    return topo->refiner.setScheme(scheme) ? 0 : -1;
}

void osdfar_topology_destroy(osdfar_topology *topo) { delete topo; }

int osdfar_topology_cache_stencils(osdfar_topology *topo, int level)
{
    try {
        if (!topo || level < 0 || level > 10 || topo->refiner.maxLevel() > level) return -1;
        topo->refiner.refineUniform(level);
        computeLevelStencils(topo->refiner, INTERPOLATE_VERTEX, topo->levels);
        return 0;
    } catch (...) {
        return OSDFAR_ERR_HOST;
    }
}

int osdfar_topology_drop_stencils(osdfar_topology *topo)
{
    if (!topo) return -1;
    std::vector<LevelStencils>().swap(topo->levels);
    return 0;
}

int osdfar_refine_uniform(osdfar_topology *topo, int maxLevel)
{
    try {
        if (!topo || maxLevel < 0 || maxLevel > 10) return -1;
        if (topo->refiner.maxLevel() > maxLevel) return -1;
        topo->refiner.refineUniform(maxLevel);
        return 0;
    } catch (...) {
        return OSDFAR_ERR_HOST;
    }
}

int osdfar_level_info(const osdfar_topology *topo, int level, int *nv, int *nf, int *ne)
{
    if (!topo || level < 0 || level > topo->refiner.maxLevel()) return -1;
    const Level &L = topo->refiner.level(level);
    if (nv) *nv = L.nv;
    if (nf) *nf = L.nf;
    if (ne) *ne = L.ne;
    return 0;
}

int osdfar_level_face_vertices(const osdfar_topology *topo, int level, int *out)
{
    if (!topo || !out || level < 0 || level > topo->refiner.maxLevel()) return -1;
    const Level &L = topo->refiner.level(level);
    std::memcpy(out, L.fv.data(), L.fv.size() * sizeof(int));
    return 0;
}

osdfar_stencil_table *osdfar_stencil_table_create(const osdfar_topology *topo, int mode,
                                                  int intermediateLevels, int controlVerts,
                                                  int useDouble)
{
    try {
        if (!topo || (mode != 0 && mode != 1)) return nullptr;
        std::unique_ptr<osdfar_stencil_table> t(new (std::nothrow) osdfar_stencil_table);
        if (!t) return nullptr;
        StencilOptions o;
        o.interpolationMode = mode;
        o.generateIntermediateLevels = intermediateLevels != 0;
        o.generateControlVerts = controlVerts != 0;
        t->isDouble = useDouble != 0;
        const bool cached = mode == 0 && (int)topo->levels.size() == topo->refiner.maxLevel() + 1;
        if (cached) {
            if (t->isDouble) stencilTableFromLevels(topo->refiner, topo->levels, o, t->d);
            else stencilTableFromLevels(topo->refiner, topo->levels, o, t->f);
        } else {
            if (t->isDouble) buildStencilTable(topo->refiner, o, t->d);
            else buildStencilTable(topo->refiner, o, t->f);
        }
        return t.release();
    } catch (...) {
        return nullptr;
    }
}

int osdfar_stencil_table_data(const osdfar_stencil_table *t, const int **sizes,
                              const int **indices, const float **weights)
{
    if (!t || t->isDouble) return -1;
    if (sizes) *sizes = t->f.sizes.data();
    if (indices) *indices = t->f.indices.data();
    if (weights) *weights = t->f.weights.data();
    return 0;
}

int osdfar_stencil_table_info(const osdfar_stencil_table *t, int *numStencils, long long *nnz,
                              int *numControlVertices, int *numWeightArrays)
{
    if (!t) return -1;
    if (t->isDouble) {
        if (numStencils) *numStencils = t->d.numStencils();
        if (nnz) *nnz = (long long)t->d.nnz();
        if (numControlVertices) *numControlVertices = t->d.numControlVertices;
    } else {
        if (numStencils) *numStencils = t->f.numStencils();
        if (nnz) *nnz = (long long)t->f.nnz();
        if (numControlVertices) *numControlVertices = t->f.numControlVertices;
    }
    if (numWeightArrays) *numWeightArrays = t->numWeightArrays;
    return 0;
}

int osdfar_stencil_table_copy(const osdfar_stencil_table *t, int *sizes, int *offsets32,
                              long long *offsets64, int *indices, int which,
                              float *weights, double *weights64)
{
    if (!t) return -1;
    return t->isDouble ? copyTable(t->d, sizes, offsets32, offsets64, indices, which, weights, weights64)
                       : copyTable(t->f, sizes, offsets32, offsets64, indices, which, weights, weights64);
}

void osdfar_stencil_table_destroy(osdfar_stencil_table *t) { delete t; }

osdfar_patch_table *osdfar_patch_table_create(osdfar_topology *topo, int isoLevel)
{
    return osdfar_patch_table_create_ex(topo, isoLevel, nullptr);
}

osdfar_patch_table *osdfar_patch_table_create_ex(osdfar_topology *topo, int isoLevel,
                                                 const osdfar_patch_options *options)
{
    try {
        if (!topo || isoLevel < 0 || isoLevel > 10) return nullptr;
        if (topo->refiner.maxLevel() > isoLevel) return nullptr;
        // Catmark and bilinear patch tables; face-varying channels for Catmark
        if (topo->refiner.scheme == SCHEME_LOOP) return nullptr;
        if (topo->refiner.scheme != SCHEME_CATMARK && topo->fvarValues > 0) return nullptr;
        PatchOptions opt;
        if (options) {
            if (options->secondaryLevel < 0 || options->secondaryLevel > 15) return nullptr;
            opt.secondaryLevel = options->secondaryLevel;
            opt.legacySharpCorners = options->legacySharpCorners != 0;
            opt.shareEndCapPoints = options->shareEndCapPoints != 0;
        }
        std::unique_ptr<osdfar_patch_table> t(new (std::nothrow) osdfar_patch_table);
        const bool cached = (int)topo->levels.size() == topo->refiner.maxLevel() + 1 &&
                            topo->refiner.maxLevel() == isoLevel;
        if (!t || !t->table.build(topo->refiner, isoLevel, cached ? &topo->levels : nullptr, opt))
            return nullptr;
        if (topo->fvarValues > 0 &&
            !t->table.buildFaceVarying(topo->refiner, topo->fvarValues, topo->fvarIndices.data()))
            return nullptr;
        return t.release();
    } catch (...) {
        return nullptr;
    }
}

void osdfar_patch_table_destroy(osdfar_patch_table *t) { delete t; }

int osdfar_patch_table_info(const osdfar_patch_table *t, int *numControl, int *numRefined,
                            int *numLocal, int *numArrays, int *numIndices, int *numPatches,
                            int *maxDepth)
{
    if (!t) return -1;
    const PatchTable &p = t->table;
    if (numControl) *numControl = p.numControl;
    if (numRefined) *numRefined = p.numRefined;
    if (numLocal) *numLocal = p.numLocal;
    if (numArrays) *numArrays = (int)p.arrays.size();
    if (numIndices) *numIndices = (int)p.indices.size();
    if (numPatches) *numPatches = (int)p.params.size();
    if (maxDepth) *maxDepth = p.maxDepth;
    return 0;
}

int osdfar_patch_table_copy(const osdfar_patch_table *t, int *arrays, int *indices, void *params)
{
    if (!t) return -1;
    const PatchTable &p = t->table;
    if (arrays) std::memcpy(arrays, p.arrays.data(), p.arrays.size() * sizeof(PatchArrayRec));
    if (indices) std::memcpy(indices, p.indices.data(), p.indices.size() * sizeof(int));
    if (params) std::memcpy(params, p.params.data(), p.params.size() * sizeof(PatchParamRec));
    return 0;
}

int osdfar_topology_set_fvar(osdfar_topology *topo, int numValues, const int *valueIndices)
{
    try {
        if (!topo || numValues < 0 || (numValues > 0 && !valueIndices)) return -1;
        const Level &L0 = topo->refiner.level(0);
        for (size_t k = 0; numValues > 0 && k < L0.fv.size(); ++k)
            if (valueIndices[k] < 0 || valueIndices[k] >= numValues) return -1;
        topo->fvarValues = numValues;
        if (numValues > 0) topo->fvarIndices.assign(valueIndices, valueIndices + L0.fv.size());
        else topo->fvarIndices.clear();
        return 0;
    } catch (...) {
        return -3;
    }
}

int osdfar_patch_table_fvar_info(const osdfar_patch_table *t, int channel, int *numBase,
                                 int *numRefined, int *numArrays, int *numIndices)
{
    if (!t || channel < 0 || channel >= (int)t->table.fvar.size()) return -1;
    const PatchTable::FVarChannel &c = t->table.fvar[channel];
    if (numBase) *numBase = c.numBase;
    if (numRefined) *numRefined = c.numRefined;
    if (numArrays) *numArrays = (int)c.arrays.size();
    if (numIndices) *numIndices = (int)c.indices.size();
    return 0;
}

int osdfar_patch_table_fvar_copy(const osdfar_patch_table *t, int channel, int *arrays,
                                 int *indices, void *params)
{
    if (!t || channel < 0 || channel >= (int)t->table.fvar.size()) return -1;
    const PatchTable::FVarChannel &c = t->table.fvar[channel];
    if (arrays) std::memcpy(arrays, c.arrays.data(), c.arrays.size() * sizeof(PatchArrayRec));
    if (indices) std::memcpy(indices, c.indices.data(), c.indices.size() * sizeof(int));
    if (params) std::memcpy(params, c.params.data(), c.params.size() * sizeof(PatchParamRec));
    return 0;
}

osdfar_stencil_table *osdfar_patch_table_fvar_stencils(const osdfar_patch_table *t, int channel,
                                                       int useDouble)
{
    try {
        if (!t || channel < 0 || channel >= (int)t->table.fvar.size()) return nullptr;
        std::unique_ptr<osdfar_stencil_table> s(new (std::nothrow) osdfar_stencil_table);
        if (!s) return nullptr;
        s->isDouble = useDouble != 0;
        if (s->isDouble) t->table.fvarStencilTable(channel, s->d);
        else t->table.fvarStencilTable(channel, s->f);
        return s.release();
    } catch (...) {
        return nullptr;
    }
}

osdfar_stencil_table *osdfar_patch_table_point_stencils(const osdfar_patch_table *t, int useDouble)
{
    try {
        if (!t) return nullptr;
        std::unique_ptr<osdfar_stencil_table> s(new (std::nothrow) osdfar_stencil_table);
        if (!s) return nullptr;
        s->isDouble = useDouble != 0;
        if (s->isDouble) t->table.pointStencilTable(s->d);
        else t->table.pointStencilTable(s->f);
        return s.release();
    } catch (...) {
        return nullptr;
    }
}

int osdfar_patch_map_find(const osdfar_patch_table *t, int n, const int *ptex, const double *u,
                          const double *v, void *patchCoords)
{
    try {
        if (!t || n < 0 || (n > 0 && (!ptex || !u || !v || !patchCoords))) return -1;
        struct Coord { int arrayIndex, patchIndex, vertIndex; float s, t; };
        Coord *out = (Coord *)patchCoords;
        int found = 0;
    #pragma omp parallel for reduction(+ : found)
        for (int i = 0; i < n; ++i) {
            PatchHandleRec h;
            if (t->table.findPatch(ptex[i], u[i], v[i], h)) {
                out[i] = { h.arrayIndex, h.patchIndex, h.vertIndex, (float)u[i], (float)v[i] };
                ++found;
            } else {
                out[i] = { -1, -1, -1, (float)u[i], (float)v[i] };
            }
        }
        return found;
    } catch (...) {
        return OSDFAR_ERR_HOST;
    }
}

osdfar_stencil_table *osdfar_limit_stencil_table_create(const osdfar_patch_table *t, int n,
                                                        const int *ptex, const double *u,
                                                        const double *v, int secondDerivatives,
                                                        int useDouble)
{
    try {
        if (!t || n < 0 || (n > 0 && (!ptex || !u || !v))) return nullptr;
        std::unique_ptr<osdfar_stencil_table> s(new (std::nothrow) osdfar_stencil_table);
        if (!s) return nullptr;
        s->isDouble = useDouble != 0;
        s->numWeightArrays = secondDerivatives ? 6 : 3;
        if (s->isDouble) t->table.limitStencils(n, ptex, u, v, secondDerivatives != 0, s->d);
        else t->table.limitStencils(n, ptex, u, v, secondDerivatives != 0, s->f);
        return s.release();
    } catch (...) {
        return nullptr;
    }
}

int osdfar_vertex_locations(const osdfar_topology *topo, int level, int *ptex, double *u,
                            double *v)
{
    try {
        if (!topo || level < 0 || level > topo->refiner.maxLevel() || !ptex || !u || !v) return -1;
        if (topo->refiner.scheme == SCHEME_LOOP) return -1;   // quad parameterizations only
        std::vector<int> p;
        std::vector<double> uu, vv;
        uniformVertexLocations(topo->refiner, level, p, uu, vv);
        std::memcpy(ptex, p.data(), p.size() * sizeof(int));
        std::memcpy(u, uu.data(), uu.size() * sizeof(double));
        std::memcpy(v, vv.data(), vv.size() * sizeof(double));
        return 0;
    } catch (...) {
        return OSDFAR_ERR_HOST;
    }
}

}  // extern "C"
