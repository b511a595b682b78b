// Thin C wrapper over the REAL OpenSubdiv reference (Far + Osd CPU/OMP
// evaluators).  TEST INFRASTRUCTURE ONLY: oracle/Makefile compiles this file
// together with the reference's own sources (read in place from the
// reference tree) into oracle/_ref/libosdref.so.  Only tests/, the golden
// fixture script and the benchmark's CPU-baseline leg load that library; the
// shipped evaluator never does.
//
// Reference APIs wrapped (paths relative to the OpenSubdiv source tree):
//   far/topologyDescriptor.h:40-80   TopologyRefinerFactory<TopologyDescriptor>
//   far/stencilTableFactory.h:55-130 StencilTableFactoryReal<REAL>::Create
//   far/stencilTableFactory.h:200-240 LimitStencilTableFactoryReal<REAL>::Create
//   far/patchTableFactory.h:40-160   PatchTableFactory::Create + options
//   osd/cpuPatchTable.h:35-100       CpuPatchTable flattening
//   far/patchMap.h:58-74             PatchMap::FindPatch
//   far/patchTable.h:429-445         PatchTable::EvaluateBasis<double>
//   osd/cpuEvaluator.h, osd/ompEvaluator.h   raw EvalStencils / EvalPatches
//   bfr/refinerSurfaceFactory.h, bfr/surface.h  Bfr::Surface per-face evaluation
#include <opensubdiv/far/topologyDescriptor.h>
#include <opensubdiv/far/topologyRefinerFactory.h>
#include <opensubdiv/far/stencilTableFactory.h>
#include <opensubdiv/far/patchTableFactory.h>
#include <opensubdiv/far/patchMap.h>
#include <opensubdiv/far/primvarRefiner.h>
#include <opensubdiv/osd/cpuEvaluator.h>
#include <opensubdiv/osd/cpuPatchTable.h>
#include <opensubdiv/osd/ompEvaluator.h>
#include <opensubdiv/bfr/tessellation.h>
#include <opensubdiv/bfr/refinerSurfaceFactory.h>
#include <opensubdiv/bfr/surface.h>

#include <cstring>
#include <vector>
#include <omp.h>

using namespace OpenSubdiv;

namespace {

struct Topo {
    int nverts = 0;
    std::vector<int> nvpf, fv, creasePairs, cornerVerts;
    std::vector<float> creaseW, cornerW;
    int boundary = 1;   // Sdc::Options::VTX_BOUNDARY_EDGE_ONLY
    int creasing = 0;   // Sdc::Options::CREASE_UNIFORM
    int scheme = Sdc::SCHEME_CATMARK;
    // one optional face-varying channel: value index per face-vertex
    int fvarValues = 0;
    std::vector<int> fvarIdx;
    int fvarInterp = Sdc::Options::FVAR_LINEAR_ALL;

    Far::TopologyRefiner *makeRefiner() const {
        Far::TopologyDescriptor d;
        d.numVertices = nverts;
        d.numFaces = (int)nvpf.size();
        d.numVertsPerFace = nvpf.data();
        d.vertIndicesPerFace = fv.data();
        d.numCreases = (int)creaseW.size();
        d.creaseVertexIndexPairs = creasePairs.empty() ? nullptr : creasePairs.data();
        d.creaseWeights = creaseW.empty() ? nullptr : creaseW.data();
        d.numCorners = (int)cornerW.size();
        d.cornerVertexIndices = cornerVerts.empty() ? nullptr : cornerVerts.data();
        d.cornerWeights = cornerW.empty() ? nullptr : cornerW.data();
        Far::TopologyDescriptor::FVarChannel ch;
        if (fvarValues > 0) {
            ch.numValues = fvarValues;
            ch.valueIndices = fvarIdx.data();
            d.numFVarChannels = 1;
            d.fvarChannels = &ch;
        }
        Sdc::Options o;
        o.SetFVarLinearInterpolation((Sdc::Options::FVarLinearInterpolation)fvarInterp);
        o.SetVtxBoundaryInterpolation((Sdc::Options::VtxBoundaryInterpolation)boundary);
        o.SetCreasingMethod((Sdc::Options::CreasingMethod)creasing);
        typedef Far::TopologyRefinerFactory<Far::TopologyDescriptor> F;
        return F::Create(d, F::Options((Sdc::SchemeType)scheme, o));
    }
};

struct Stencils {
    int nControl = 0;
    std::vector<int> sizes, offsets, indices;
    std::vector<double> w[6];   // P, du, dv, duu, duv, dvv (as double)
    std::vector<float> wf[6];   // the float table exactly as Far made it (if float)
    bool isDouble = false;

    size_t nnz = 0;
    // Far's weight vectors may carry unused capacity past the last stencil;
    // every array is cut to sum(sizes) entries.
    template <class T> void fromTable(T const *t) {
        nControl = t->GetNumControlVertices();
        sizes = t->GetSizes();
        offsets = t->GetOffsets();
        nnz = 0;
        for (int s : sizes) nnz += (size_t)s;
        indices.assign(t->GetControlIndices().begin(), t->GetControlIndices().begin() + nnz);
        setW(0, t->GetWeights());
    }
    template <class V> void setW(int k, V const &v) {
        w[k].assign(v.begin(), v.begin() + nnz);
        if (!isDouble) wf[k].assign(v.begin(), v.begin() + nnz);
    }
};

struct Patches {
    Far::TopologyRefiner *refiner = nullptr;
    Far::PatchTable const *table = nullptr;
    Far::PatchMap *map = nullptr;
    Osd::CpuPatchTable *flat = nullptr;
    Stencils st;   // coarse -> (refined + local points)
    Stencils fst;  // face-varying channel 0: base values -> (refined + local values)
    int nCoarse = 0;
    int nFVarBase = 0;
    ~Patches() { delete map; delete flat; delete table; delete refiner; }
};

template <class TBL> void fillStencils(Stencils &s, TBL const *t) { s.fromTable(t); }

}  // namespace

extern "C" {

void *ref_topo_create(int nverts, int nfaces, const int *nvpf, const int *fv,
                      int ncreases, const int *creasePairs, const float *creaseW,
                      int ncorners, const int *cornerVerts, const float *cornerW,
                      int boundary, int creasing)
{
    Topo *t = new Topo;
    t->nverts = nverts;
    t->nvpf.assign(nvpf, nvpf + nfaces);
    int nfv = 0;
    for (int i = 0; i < nfaces; ++i) nfv += nvpf[i];
    t->fv.assign(fv, fv + nfv);
    if (ncreases > 0) {
        t->creasePairs.assign(creasePairs, creasePairs + 2 * ncreases);
        t->creaseW.assign(creaseW, creaseW + ncreases);
    }
    if (ncorners > 0) {
        t->cornerVerts.assign(cornerVerts, cornerVerts + ncorners);
        t->cornerW.assign(cornerW, cornerW + ncorners);
    }
    t->boundary = boundary;
    t->creasing = creasing;
    return t;
}

void ref_topo_destroy(void *t) { delete (Topo *)t; }

// Subdivision scheme of later refiners: Sdc::SchemeType (0 bilinear,
// 1 Catmark, 2 Loop; sdc/types.h:36-40).
void ref_topo_set_scheme(void *t, int scheme) { ((Topo *)t)->scheme = scheme; }

// One face-varying channel on later refiners (far/topologyDescriptor.h:
// 73-84): numValues values, one value index per face-vertex, and
// Sdc::Options::FVarLinearInterpolation (sdc/options.h:60-67).
void ref_topo_set_fvar(void *t, int numValues, const int *valueIndices, int linearInterp)
{
    Topo *p = (Topo *)t;
    p->fvarValues = numValues;
    p->fvarIdx.assign(valueIndices, valueIndices + p->fv.size());
    p->fvarInterp = linearInterp;
}

// Uniform refinement to `level`, then StencilTableFactory::Create.
void *ref_stencils_uniform(void *topo, int level, int mode, int genIntermediate,
                           int factorize, int genControl, int useDouble)
{
    Far::TopologyRefiner *r = ((Topo *)topo)->makeRefiner();
    if (!r) return nullptr;
    r->RefineUniform(Far::TopologyRefiner::UniformOptions(level));
    Stencils *s = new Stencils;
    s->isDouble = useDouble != 0;
    if (useDouble) {
        Far::StencilTableFactoryReal<double>::Options o;
        o.interpolationMode = mode;
        o.generateOffsets = true;
        o.generateIntermediateLevels = genIntermediate;
        o.factorizeIntermediateLevels = factorize;
        o.generateControlVerts = genControl;
        auto *t = Far::StencilTableFactoryReal<double>::Create(*r, o);
        fillStencils(*s, t);
        delete t;
    } else {
        Far::StencilTableFactory::Options o;
        o.interpolationMode = mode;
        o.generateOffsets = true;
        o.generateIntermediateLevels = genIntermediate;
        o.factorizeIntermediateLevels = factorize;
        o.generateControlVerts = genControl;
        auto *t = Far::StencilTableFactory::Create(*r, o);
        fillStencils(*s, t);
        delete t;
    }
    delete r;
    return s;
}

// LimitStencilTableFactory on an adaptively refined refiner (isolation
// `isoLevel`), one location per (ptex, s, t) triple.  Locations of the same
// ptex face that are adjacent in the input are grouped into one LocationArray.
void *ref_stencils_limit(void *topo, int isoLevel, int n, const int *ptex,
                         const double *s, const double *t, int d2, int useDouble)
{
    Far::TopologyRefiner *r = ((Topo *)topo)->makeRefiner();
    if (!r) return nullptr;
    Far::TopologyRefiner::AdaptiveOptions ao(isoLevel);
    r->RefineAdaptive(ao);
    Stencils *st = new Stencils;
    st->isDouble = useDouble != 0;
    auto run = [&](auto tag) {
        typedef decltype(tag) REAL;
        std::vector<REAL> ss(s, s + n), tt(t, t + n);
        typename Far::LimitStencilTableFactoryReal<REAL>::LocationArrayVec locs;
        for (int i = 0; i < n;) {
            int j = i;
            while (j < n && ptex[j] == ptex[i]) ++j;
            typename Far::LimitStencilTableFactoryReal<REAL>::LocationArray a;
            a.ptexIdx = ptex[i];
            a.numLocations = j - i;
            a.s = &ss[i];
            a.t = &tt[i];
            locs.push_back(a);
            i = j;
        }
        typename Far::LimitStencilTableFactoryReal<REAL>::Options o;
        o.generate1stDerivatives = true;
        o.generate2ndDerivatives = d2 != 0;
        auto *lt = Far::LimitStencilTableFactoryReal<REAL>::Create(*r, locs, nullptr, nullptr, o);
        st->fromTable(lt);
        st->setW(1, lt->GetDuWeights());
        st->setW(2, lt->GetDvWeights());
        if (d2) {
            st->setW(3, lt->GetDuuWeights());
            st->setW(4, lt->GetDuvWeights());
            st->setW(5, lt->GetDvvWeights());
        }
        delete lt;
    };
    if (useDouble) run(double()); else run(float());
    delete r;
    return st;
}

void ref_stencils_info(void *h, int *nStencils, int *nnz, int *nControl)
{
    Stencils *s = (Stencils *)h;
    *nStencils = (int)s->sizes.size();
    *nnz = (int)s->indices.size();
    *nControl = s->nControl;
}

// which: 0=P 1=du 2=dv 3=duu 4=duv 5=dvv.  Pointers may be NULL.
void ref_stencils_copy(void *h, int *sizes, int *offsets, int *indices,
                       int which, float *wf, double *wd)
{
    Stencils *s = (Stencils *)h;
    if (sizes) std::memcpy(sizes, s->sizes.data(), s->sizes.size() * sizeof(int));
    if (offsets) std::memcpy(offsets, s->offsets.data(), s->offsets.size() * sizeof(int));
    if (indices) std::memcpy(indices, s->indices.data(), s->indices.size() * sizeof(int));
    if (wf) {
        if (!s->wf[which].empty())
            std::memcpy(wf, s->wf[which].data(), s->wf[which].size() * sizeof(float));
        else
            for (size_t i = 0; i < s->w[which].size(); ++i) wf[i] = (float)s->w[which][i];
    }
    if (wd) std::memcpy(wd, s->w[which].data(), s->w[which].size() * sizeof(double));
}

void ref_stencils_destroy(void *h) { delete (Stencils *)h; }

// Feature-adaptive patch table as in examples/glEvalLimit (glEvalLimit.cpp:
// 670-760): RefineAdaptive(iso), vertex stencils with intermediate levels,
// PatchTableFactory, AppendLocalPointStencilTable, CpuPatchTable, PatchMap.
void *ref_patches_create(void *topo, int isoLevel, int endcap, int infSharp,
                         int singleCrease, int legacySharpCorner, int useDouble)
{
    Patches *p = new Patches;
    p->refiner = ((Topo *)topo)->makeRefiner();
    if (!p->refiner) { delete p; return nullptr; }
    p->nCoarse = p->refiner->GetLevel(0).GetNumVertices();
    Far::PatchTableFactory::Options po(isoLevel);
    po.SetEndCapType((Far::PatchTableFactory::Options::EndCapType)endcap);
    po.useInfSharpPatch = infSharp;
    po.useSingleCreasePatch = singleCrease;
    po.generateLegacySharpCornerPatches = legacySharpCorner;
    po.generateVaryingTables = false;
    if (useDouble) po.SetPatchPrecision<double>();
    p->refiner->RefineAdaptive(po.GetRefineAdaptiveOptions());
    p->table = Far::PatchTableFactory::Create(*p->refiner, po);
    p->st.isDouble = useDouble != 0;
    auto build = [&](auto tag) {
        typedef decltype(tag) REAL;
        typename Far::StencilTableFactoryReal<REAL>::Options so;
        so.generateOffsets = true;
        so.generateIntermediateLevels = true;
        auto *base = Far::StencilTableFactoryReal<REAL>::Create(*p->refiner, so);
        auto const *local = p->table->template GetLocalPointStencilTable<REAL>();
        auto const *all = base;
        if (local) {
            all = Far::StencilTableFactoryReal<REAL>::AppendLocalPointStencilTable(*p->refiner, base, local);
            delete base;
        }
        p->st.fromTable(all);
        delete all;
    };
    if (useDouble) build(double()); else build(float());
    p->flat = new Osd::CpuPatchTable(p->table);
    p->map = new Far::PatchMap(*p->table);
    return p;
}

// Feature-adaptive patch table with Far's secondary isolation level
// (TopologyRefiner::AdaptiveOptions::secondaryLevel, far/topologyRefiner.h:
// 156-182): features other than smooth extraordinary vertices are isolated
// to isoLevel, smooth extraordinary vertices only to secondaryLevel.
// Gregory end caps, end-cap points shared or not, legacy sharp-corner
// patches on or off -- the set-up regression/bfr_evaluate/
// farPatchEvaluator.cpp:52-84 compares Bfr with (useInfSharpPatch as
// given).
void *ref_patches_create_ex(void *topo, int isoLevel, int secondaryLevel, int infSharp,
                            int legacySharpCorner, int shareEndCapPoints, int useDouble)
{
    Patches *p = new Patches;
    p->refiner = ((Topo *)topo)->makeRefiner();
    if (!p->refiner) { delete p; return nullptr; }
    p->nCoarse = p->refiner->GetLevel(0).GetNumVertices();
    Far::PatchTableFactory::Options po(isoLevel);
    po.SetEndCapType(Far::PatchTableFactory::Options::ENDCAP_GREGORY_BASIS);
    po.useInfSharpPatch = infSharp != 0;
    po.generateLegacySharpCornerPatches = legacySharpCorner != 0;
    po.shareEndCapPatchPoints = shareEndCapPoints != 0;
    po.generateVaryingTables = false;
    if (useDouble) po.SetPatchPrecision<double>();
    Far::TopologyRefiner::AdaptiveOptions ro = po.GetRefineAdaptiveOptions();
    ro.SetIsolationLevel(isoLevel);
    ro.SetSecondaryLevel(secondaryLevel);
    p->refiner->RefineAdaptive(ro);
    p->table = Far::PatchTableFactory::Create(*p->refiner, po);
    p->st.isDouble = useDouble != 0;
    auto build = [&](auto tag) {
        typedef decltype(tag) REAL;
        typename Far::StencilTableFactoryReal<REAL>::Options so;
        so.generateOffsets = true;
        so.generateIntermediateLevels = true;
        auto *base = Far::StencilTableFactoryReal<REAL>::Create(*p->refiner, so);
        auto const *local = p->table->template GetLocalPointStencilTable<REAL>();
        auto const *all = base;
        if (local) {
            all = Far::StencilTableFactoryReal<REAL>::AppendLocalPointStencilTable(*p->refiner, base, local);
            delete base;
        }
        p->st.fromTable(all);
        delete all;
    };
    if (useDouble) build(double()); else build(float());
    p->flat = new Osd::CpuPatchTable(p->table);
    p->map = new Far::PatchMap(*p->table);
    return p;
}

// Bfr::Surface<double> of base face faces[i] evaluated at (u[i], v[i]) in
// the face's parameterization (bfr/surface.h:227-266; the set-up of
// tutorials/bfr/tutorial_1_1): RefinerSurfaceFactory over the unrefined
// topology with approxLevelSmooth / approxLevelSharp, InitVertexSurface,
// PreparePatchPoints from the 3-float control points, then Evaluate
// (P, Du, Dv; 3 doubles each per query).  Returns the number of queries
// whose face had a valid surface (others are left untouched).
int ref_bfr_evaluate(void *topo, int approxSmooth, int approxSharp, const double *pos, int n,
                     const int *faces, const double *u, const double *v, double *P, double *Du,
                     double *Dv)
{
    Far::TopologyRefiner *r = ((Topo *)topo)->makeRefiner();
    if (!r) return -1;
    int done = 0;
    {
        Bfr::SurfaceFactory::Options so;
        so.SetApproxLevelSmooth(approxSmooth);
        so.SetApproxLevelSharp(approxSharp);
        Bfr::RefinerSurfaceFactory<> factory(*r, so);
        Bfr::Surface<double> surface;
        std::vector<double> pts;
        int current = -1;
        bool valid = false;
        for (int i = 0; i < n; ++i) {
            if (faces[i] != current) {
                current = faces[i];
                valid = factory.InitVertexSurface(current, &surface);
                if (valid) {
                    pts.resize((size_t)surface.GetNumPatchPoints() * 3);
                    surface.PreparePatchPoints(pos, 3, pts.data(), 3);
                }
            }
            if (!valid) continue;
            const double uv[2] = { u[i], v[i] };
            surface.Evaluate(uv, pts.data(), 3, P + 3 * (size_t)i, Du + 3 * (size_t)i,
                             Dv + 3 * (size_t)i);
            ++done;
        }
    }
    delete r;
    return done;
}

// Uniform linear patch table (far/patchTableFactory.cpp:745-850,
// BuildUniformPolygons): RefineUniform(level), one QUADS or TRIANGLES patch
// per face of the last level (TRIANGLES for Loop meshes, or Catmark quads
// split in two with triangulateQuads), indices into [coarse | refined
// levels], the stencil table with intermediate levels that produces them,
// and the table's PatchMap.
void *ref_patches_create_uniform(void *topo, int level, int triangulateQuads)
{
    Patches *p = new Patches;
    p->refiner = ((Topo *)topo)->makeRefiner();
    if (!p->refiner) { delete p; return nullptr; }
    p->nCoarse = p->refiner->GetLevel(0).GetNumVertices();
    p->refiner->RefineUniform(Far::TopologyRefiner::UniformOptions(level));
    Far::PatchTableFactory::Options po(level);
    po.triangulateQuads = triangulateQuads != 0;
    po.generateVaryingTables = false;
    p->table = Far::PatchTableFactory::Create(*p->refiner, po);
    Far::StencilTableFactory::Options so;
    so.generateOffsets = true;
    so.generateIntermediateLevels = true;
    auto const *st = Far::StencilTableFactory::Create(*p->refiner, so);
    p->st.fromTable(st);
    delete st;
    p->flat = new Osd::CpuPatchTable(p->table);
    p->map = new Far::PatchMap(*p->table);
    return p;
}

// Feature-adaptive patch table with the topology's face-varying channel,
// set up as examples/glEvalLimit/glEvalLimit.cpp:684-750 does: vertex and
// face-varying (INTERPOLATE_FACE_VARYING, channel 0) stencil tables with
// intermediate levels, PatchTableFactory with generateFVarTables (and
// generateFVarLegacyLinearPatches as given: Far's default is true, linear
// QUADS face-varying patches; false builds B-spline / Gregory face-varying
// patches unless the channel interpolates linearly everywhere), local points
// appended to both stencil tables.
void *ref_patches_create_fvar(void *topo, int isoLevel, int legacyLinear)
{
    Topo *tp = (Topo *)topo;
    if (tp->fvarValues <= 0) return nullptr;
    Patches *p = new Patches;
    p->refiner = tp->makeRefiner();
    if (!p->refiner) { delete p; return nullptr; }
    p->nCoarse = p->refiner->GetLevel(0).GetNumVertices();
    p->nFVarBase = p->refiner->GetLevel(0).GetNumFVarValues(0);
    Far::PatchTableFactory::Options po(isoLevel);
    po.SetEndCapType(Far::PatchTableFactory::Options::ENDCAP_GREGORY_BASIS);
    po.generateVaryingTables = false;
    po.generateFVarTables = true;
    po.generateFVarLegacyLinearPatches = legacyLinear != 0;
    p->refiner->RefineAdaptive(po.GetRefineAdaptiveOptions());
    p->table = Far::PatchTableFactory::Create(*p->refiner, po);
    Far::StencilTableFactory::Options so;
    so.generateOffsets = true;
    so.generateIntermediateLevels = true;
    Far::StencilTable const *vs = Far::StencilTableFactory::Create(*p->refiner, so);
    if (auto const *local = p->table->GetLocalPointStencilTable()) {
        auto const *all = Far::StencilTableFactory::AppendLocalPointStencilTable(*p->refiner, vs, local);
        delete vs;
        vs = all;
    }
    so.interpolationMode = Far::StencilTableFactory::INTERPOLATE_FACE_VARYING;
    so.fvarChannel = 0;
    Far::StencilTable const *fs = Far::StencilTableFactory::Create(*p->refiner, so);
    if (auto const *local = p->table->GetLocalPointFaceVaryingStencilTable(0)) {
        auto const *all = Far::StencilTableFactory::AppendLocalPointStencilTableFaceVarying(
            *p->refiner, fs, local, 0);
        delete fs;
        fs = all;
    }
    p->st.fromTable(vs);
    p->fst.fromTable(fs);
    p->fst.nControl = p->nFVarBase;
    delete vs;
    delete fs;
    p->flat = new Osd::CpuPatchTable(p->table);
    p->map = new Far::PatchMap(*p->table);
    return p;
}

// Face-varying channel 0 of a table from ref_patches_create_fvar: base value
// count, its stencil table's size, the flattened array / index / param counts
// (osd/cpuPatchTable.cpp:54-115) and the regular / irregular patch types.
void ref_patches_fvar_info(void *h, int *nBase, int *nStencils, int *nnz, int *nArrays,
                           int *nIndices, int *nParams, int *regType, int *irregType)
{
    Patches *p = (Patches *)h;
    *nBase = p->nFVarBase;
    *nStencils = (int)p->fst.sizes.size();
    *nnz = (int)p->fst.indices.size();
    *nArrays = (int)p->flat->GetNumPatchArrays();
    *nIndices = (int)p->flat->GetFVarPatchIndexSize(0);
    *nParams = (int)p->flat->GetFVarPatchParamSize(0);
    *regType = (int)p->table->GetFVarPatchDescriptorRegular(0).GetType();
    *irregType = (int)p->table->GetFVarPatchDescriptorIrregular(0).GetType();
}

void *ref_patches_fvar_stencils(void *h) { return &((Patches *)h)->fst; }

void ref_patches_fvar_copy(void *h, int *arrays, int *indices, void *params)
{
    Patches *p = (Patches *)h;
    if (arrays)
        std::memcpy(arrays, p->flat->GetFVarPatchArrayBuffer(0),
                    p->flat->GetNumPatchArrays() * sizeof(Osd::PatchArray));
    if (indices)
        std::memcpy(indices, p->flat->GetFVarPatchIndexBuffer(0),
                    p->flat->GetFVarPatchIndexSize(0) * sizeof(int));
    if (params)
        std::memcpy(params, p->flat->GetFVarPatchParamBuffer(0),
                    p->flat->GetFVarPatchParamSize(0) * sizeof(Osd::PatchParam));
}

void ref_patches_info(void *h, int *nCoarse, int *nStencils, int *nnz,
                      int *nArrays, int *nIndices, int *nParams)
{
    Patches *p = (Patches *)h;
    *nCoarse = p->nCoarse;
    *nStencils = (int)p->st.sizes.size();
    *nnz = (int)p->st.indices.size();
    *nArrays = (int)p->flat->GetNumPatchArrays();
    *nIndices = (int)p->flat->GetPatchIndexSize();
    *nParams = (int)p->flat->GetPatchParamSize();
}

void *ref_patches_stencils(void *h) { return &((Patches *)h)->st; }

// arrays: 6 ints each; params: field0, field1, sharpness (as 3 x 4 bytes).
void ref_patches_copy(void *h, int *arrays, int *indices, void *params)
{
    Patches *p = (Patches *)h;
    if (arrays)
        std::memcpy(arrays, p->flat->GetPatchArrayBuffer(),
                    p->flat->GetNumPatchArrays() * sizeof(Osd::PatchArray));
    if (indices)
        std::memcpy(indices, p->flat->GetPatchIndexBuffer(),
                    p->flat->GetPatchIndexSize() * sizeof(int));
    if (params)
        std::memcpy(params, p->flat->GetPatchParamBuffer(),
                    p->flat->GetPatchParamSize() * sizeof(Osd::PatchParam));
}

// PatchMap::FindPatch; out = 3 ints per query (arrayIndex, patchIndex,
// vertIndex), -1s when no patch is found.  Returns the number found.
int ref_patches_find(void *h, int n, const int *faces, const double *u,
                     const double *v, int *out)
{
    Patches *p = (Patches *)h;
    int found = 0;
    if (!p->map) return -1;
    for (int i = 0; i < n; ++i) {
        auto const *hd = p->map->FindPatch(faces[i], u[i], v[i]);
        if (hd) {
            out[3 * i] = hd->arrayIndex;
            out[3 * i + 1] = hd->patchIndex;
            out[3 * i + 2] = hd->vertIndex;
            ++found;
        } else {
            out[3 * i] = out[3 * i + 1] = out[3 * i + 2] = -1;
        }
    }
    return found;
}

// Far::PatchTable::EvaluateBasis<double>; weights are 20 per query
// (unused tail left at 0); returns per-query point count in npts.
void ref_patches_eval_basis_f64(void *h, int n, const int *handles,
                                const double *s, const double *t,
                                double *wP, double *wDs, double *wDt, int *npts)
{
    Patches *p = (Patches *)h;
    for (int i = 0; i < n; ++i) {
        Far::PatchTable::PatchHandle hd;
        hd.arrayIndex = handles[3 * i];
        hd.patchIndex = handles[3 * i + 1];
        hd.vertIndex = handles[3 * i + 2];
        double *a = wP + 20 * i, *b = wDs + 20 * i, *c = wDt + 20 * i;
        std::memset(a, 0, 20 * sizeof(double));
        std::memset(b, 0, 20 * sizeof(double));
        std::memset(c, 0, 20 * sizeof(double));
        p->table->EvaluateBasis(hd, s[i], t[i], a, b, c);
        npts[i] = p->table->GetPatchDescriptor(hd).GetNumControlVertices();
    }
}

void ref_patches_destroy(void *h) { delete (Patches *)h; }

// Uniform refinement of a primvar through PrimvarRefiner + Limit at the
// last level (position and two tangents) -- an independent path to limit
// positions of vertices (far/primvarRefiner.h:164-200).
struct V3 {
    double x[3];
    void Clear(void * = 0) { x[0] = x[1] = x[2] = 0; }
    void AddWithWeight(V3 const &s, double w) { for (int k = 0; k < 3; ++k) x[k] += w * s.x[k]; }
};

int ref_limit_points_uniform(void *topo, int level, const double *coarse,
                             double *pos, double *tan1, double *tan2)
{
    Far::TopologyRefiner *r = ((Topo *)topo)->makeRefiner();
    Far::TopologyRefiner::UniformOptions uo(level);
    uo.fullTopologyInLastLevel = true;      // required by PrimvarRefiner::Limit
    r->RefineUniform(uo);
    int total = r->GetNumVerticesTotal();
    std::vector<V3> v(total);
    std::memcpy(v.data(), coarse, sizeof(double) * 3 * r->GetLevel(0).GetNumVertices());
    Far::PrimvarRefinerReal<double> pr(*r);
    V3 *src = v.data();
    for (int l = 1; l <= level; ++l) {
        V3 *dst = src + r->GetLevel(l - 1).GetNumVertices();
        pr.Interpolate(l, src, dst);
        src = dst;
    }
    int nLast = r->GetLevel(level).GetNumVertices();
    std::vector<V3> P(nLast), T1(nLast), T2(nLast);
    pr.Limit(src, P, T1, T2);
    std::memcpy(pos, P.data(), sizeof(double) * 3 * nLast);
    if (tan1) std::memcpy(tan1, T1.data(), sizeof(double) * 3 * nLast);
    if (tan2) std::memcpy(tan2, T2.data(), sizeof(double) * 3 * nLast);
    delete r;
    return nLast;
}

// ---- Osd::CpuEvaluator / Osd::OmpEvaluator raw entry points -------------

static Osd::BufferDescriptor D(const int *d) { return Osd::BufferDescriptor(d[0], d[1], d[2]); }

int ref_cpu_eval_stencils(const float *src, const int *sd, float *dst, const int *dd,
                          const int *sizes, const int *offsets, const int *indices,
                          const float *weights, int start, int end)
{
    return Osd::CpuEvaluator::EvalStencils(src, D(sd), dst, D(dd), sizes, offsets,
                                           indices, weights, start, end);
}

int ref_cpu_eval_stencils_d1(const float *src, const int *sd, float *dst, const int *dd,
                             float *du, const int *dud, float *dv, const int *dvd,
                             const int *sizes, const int *offsets, const int *indices,
                             const float *w, const float *duw, const float *dvw,
                             int start, int end)
{
    return Osd::CpuEvaluator::EvalStencils(src, D(sd), dst, D(dd), du, D(dud), dv, D(dvd),
                                           sizes, offsets, indices, w, duw, dvw, start, end);
}

int ref_cpu_eval_stencils_d2(const float *src, const int *sd, float *dst, const int *dd,
                             float *du, const int *dud, float *dv, const int *dvd,
                             float *duu, const int *duud, float *duv, const int *duvd,
                             float *dvv, const int *dvvd,
                             const int *sizes, const int *offsets, const int *indices,
                             const float *w, const float *duw, const float *dvw,
                             const float *duuw, const float *duvw, const float *dvvw,
                             int start, int end)
{
    return Osd::CpuEvaluator::EvalStencils(src, D(sd), dst, D(dd), du, D(dud), dv, D(dvd),
                                           duu, D(duud), duv, D(duvd), dvv, D(dvvd),
                                           sizes, offsets, indices, w, duw, dvw,
                                           duuw, duvw, dvvw, start, end);
}

int ref_omp_eval_stencils(const float *src, const int *sd, float *dst, const int *dd,
                          const int *sizes, const int *offsets, const int *indices,
                          const float *weights, int start, int end, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return Osd::OmpEvaluator::EvalStencils(src, D(sd), dst, D(dd), sizes, offsets,
                                           indices, weights, start, end);
}

int ref_omp_eval_stencils_d1(const float *src, const int *sd, float *dst, const int *dd,
                             float *du, const int *dud, float *dv, const int *dvd,
                             const int *sizes, const int *offsets, const int *indices,
                             const float *w, const float *duw, const float *dvw,
                             int start, int end, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return Osd::OmpEvaluator::EvalStencils(src, D(sd), dst, D(dd), du, D(dud), dv, D(dvd),
                                           sizes, offsets, indices, w, duw, dvw, start, end);
}

int ref_cpu_eval_patches(const float *src, const int *sd, float *dst, const int *dd,
                         float *du, const int *dud, float *dv, const int *dvd,
                         int n, const void *coords, const void *arrays,
                         const int *indices, const void *params)
{
    if (!du && !dv)
        return Osd::CpuEvaluator::EvalPatches(src, D(sd), dst, D(dd), n,
                                              (const Osd::PatchCoord *)coords,
                                              (const Osd::PatchArray *)arrays, indices,
                                              (const Osd::PatchParam *)params);
    return Osd::CpuEvaluator::EvalPatches(src, D(sd), dst, D(dd), du, D(dud), dv, D(dvd), n,
                                          (const Osd::PatchCoord *)coords,
                                          (const Osd::PatchArray *)arrays, indices,
                                          (const Osd::PatchParam *)params);
}

int ref_cpu_eval_patches_d2(const float *src, const int *sd, float *dst, const int *dd,
                            float *du, const int *dud, float *dv, const int *dvd,
                            float *duu, const int *duud, float *duv, const int *duvd,
                            float *dvv, const int *dvvd,
                            int n, const void *coords, const void *arrays,
                            const int *indices, const void *params)
{
    return Osd::CpuEvaluator::EvalPatches(src, D(sd), dst, D(dd), du, D(dud), dv, D(dvd),
                                          duu, D(duud), duv, D(duvd), dvv, D(dvvd), n,
                                          (const Osd::PatchCoord *)coords,
                                          (const Osd::PatchArray *)arrays, indices,
                                          (const Osd::PatchParam *)params);
}

int ref_omp_eval_patches(const float *src, const int *sd, float *dst, const int *dd,
                         float *du, const int *dud, float *dv, const int *dvd,
                         int n, const void *coords, const void *arrays,
                         const int *indices, const void *params, int nthreads)
{
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return Osd::OmpEvaluator::EvalPatches(src, D(sd), dst, D(dd), du, D(dud), dv, D(dvd), n,
                                          (const Osd::PatchCoord *)coords,
                                          (const Osd::PatchArray *)arrays, indices,
                                          (const Osd::PatchParam *)params);
}

// Bfr::Tessellation's coordinates (bfr/tessellation.h:51-240) of a Catmark
// face of faceSize corners at a uniform rate, float (u,v) pairs; returns
// the count, or -1 for an invalid pattern.
int ref_tess_uniform_coords(int faceSize, int rate, float *uv)
{
    OpenSubdiv::Bfr::Parameterization p(OpenSubdiv::Sdc::SCHEME_CATMARK, faceSize);
    if (!p.IsValid()) return -1;
    OpenSubdiv::Bfr::Tessellation t(p, rate);
    if (!t.IsValid()) return -1;
    return t.GetCoords(uv);
}

// Bfr::Tessellation(Parameterization(scheme, faceSize), numRates, rates)
// .GetCoords<float> (bfr/tessellation.h:107-240); uv NULL returns the
// count; -1 for an invalid parameterization or rates.
int ref_tess_coords(int scheme, int faceSize, int numRates, const int *rates, float *uv)
{
    OpenSubdiv::Bfr::Parameterization p((OpenSubdiv::Sdc::SchemeType)scheme, faceSize);
    if (!p.IsValid()) return -1;
    OpenSubdiv::Bfr::Tessellation t(p, numRates, rates);
    if (!t.IsValid()) return -1;
    return uv ? t.GetCoords(uv) : t.GetNumCoords();
}

}  // extern "C"
