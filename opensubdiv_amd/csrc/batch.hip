// Batched evaluation (osdhip_batch_*): segmented stencil and patch work of
// many meshes and frames, captured once into a HIP graph of
//   [counter reset] -> phase A (k_batch_stencils) -> phase B (k_batch_patches)
// and replayed by osdhip_batch_launch.  Tuning knobs select the kernel
// variant (frames per chunk, gather group, slices per work item, static or
// persistent scheduling); osdhip_batch_tune times the variants on the
// batch's own data and keeps the fastest.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/osdhip.h"
#include "batch_kernels.h"
#include "common.h"
#include "table.h"

using namespace osdhip;

constexpr int kCounters = kNumXcd;      // claim counters per (stage, phase)

struct BatchConfig {
    int aFrames = 2, aGroup = 4, aItemSlices = 8, aPersistent = 0;
    int bFrames = 1, bPersistent = 0, bStage = 1;
};

struct osdhip_batch {
    int length = 0;
    int nS = 0, nP = 0;
    bool tiled = false;                    // stencil tables in layout 2
    long long stencilRows = 0, frameRows = 0, patchCoords = 0, patchFrames = 0, tableBytes = 0;
    BatchStencilSeg *dS = nullptr;
    BatchPatchSeg *dP = nullptr;
    std::vector<BatchStencilSeg> hS;
    std::vector<BatchPatchSeg> hP;
    std::vector<const osdhip_stencil_table *> tables;   // per stencil segment (plain path)
    std::vector<int> sStage, pStage;       // stage of each segment
    int nStages = 1;
    BatchItem *itemsA = nullptr, *itemsB = nullptr;
    int nItemsA = 0, nItemsB = 0, itemsASlices = 0;
    std::vector<int> beginA, countA;       // phase-A item ranges per stage
    // phase-B item ranges per (launch kind, stage): each launch visits only
    // the items with waves of its kind
    std::vector<int> beginB[5], countB[5];
    std::vector<std::vector<int>> hNuniq, hNgen;   // per patch segment, while building
    unsigned *counters = nullptr;          // kCounters per (stage, phase)
    int numCUs = 256;
    BatchConfig cfg;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> prep;              // patch point staging lists
    long long stagedCoords = 0;            // coords in listed waves
};

namespace {

template <int LC, int FC, int G>
void launchA(const BatchArgs &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_batch_stencils<LC, FC, G>), dim3(grid), dim3(kBlock), 0, st, a);
}

template <int FC, int G>
bool launchA_lc(int lc, const BatchArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: launchA<1, FC, G>(a, grid, st); return true;
    case 2: launchA<2, FC, G>(a, grid, st); return true;
    case 3: launchA<3, FC, G>(a, grid, st); return true;
    case 4: launchA<4, FC, G>(a, grid, st); return true;
    case 5: launchA<5, FC, G>(a, grid, st); return true;
    case 6: launchA<6, FC, G>(a, grid, st); return true;
    case 7: launchA<7, FC, G>(a, grid, st); return true;
    case 8: launchA<8, FC, G>(a, grid, st); return true;
    default: return false;
    }
}

template <int LC>
void launchRR(const BatchArgs &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL((k_batch_stencils_rr<LC>), dim3(grid), dim3(kBlock), 0, st, a);
}

bool launchRR_lc(int lc, const BatchArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: launchRR<1>(a, grid, st); return true;
    case 2: launchRR<2>(a, grid, st); return true;
    case 3: launchRR<3>(a, grid, st); return true;
    case 4: launchRR<4>(a, grid, st); return true;
    case 5: launchRR<5>(a, grid, st); return true;
    case 6: launchRR<6>(a, grid, st); return true;
    case 7: launchRR<7>(a, grid, st); return true;
    case 8: launchRR<8>(a, grid, st); return true;
    default: return false;
    }
}

template <int FC, int G>
bool launchT_lc(int lc, const BatchArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_batch_stencils_tiled<1, FC, G>), dim3(grid), dim3(kBlock), 0, st, a); return true;
    case 2: hipLaunchKernelGGL((k_batch_stencils_tiled<2, FC, G>), dim3(grid), dim3(kBlock), 0, st, a); return true;
    case 3: hipLaunchKernelGGL((k_batch_stencils_tiled<3, FC, G>), dim3(grid), dim3(kBlock), 0, st, a); return true;
    case 4: hipLaunchKernelGGL((k_batch_stencils_tiled<4, FC, G>), dim3(grid), dim3(kBlock), 0, st, a); return true;
    case 5: hipLaunchKernelGGL((k_batch_stencils_tiled<5, FC, G>), dim3(grid), dim3(kBlock), 0, st, a); return true;
    case 6: hipLaunchKernelGGL((k_batch_stencils_tiled<6, FC, G>), dim3(grid), dim3(kBlock), 0, st, a); return true;
    case 7: hipLaunchKernelGGL((k_batch_stencils_tiled<7, FC, G>), dim3(grid), dim3(kBlock), 0, st, a); return true;
    case 8: hipLaunchKernelGGL((k_batch_stencils_tiled<8, FC, G>), dim3(grid), dim3(kBlock), 0, st, a); return true;
    default: return false;
    }
}

// phase A variants on locality-ordered tables: (frames per chunk, group)
bool launchT_variant(int lc, int fc, int g, const BatchArgs &a, int grid, hipStream_t st)
{
    if (fc == 1 && g == 4) return launchT_lc<1, 4>(lc, a, grid, st);
    if (fc == 1 && g == 8) return launchT_lc<1, 8>(lc, a, grid, st);
    if (fc == 2 && g == 4) return launchT_lc<2, 4>(lc, a, grid, st);
    if (fc == 2 && g == 8) return launchT_lc<2, 8>(lc, a, grid, st);
    if (fc == 4 && g == 4) return launchT_lc<4, 4>(lc, a, grid, st);
    if (fc == 4 && g == 8) return launchT_lc<4, 8>(lc, a, grid, st);
    return false;
}

// phase A variants: (frames per chunk, gather group); frames 0 = the
// records of a slice stay in registers for all frames (group ignored)
bool launchA_variant(bool tiled, int lc, int fc, int g, const BatchArgs &a, int grid,
                     hipStream_t st)
{
    if (tiled) return launchT_variant(lc, fc, g, a, grid, st);
    if (fc == 0) return launchRR_lc(lc, a, grid, st);
    if (fc == 1 && g == 4) return launchA_lc<1, 4>(lc, a, grid, st);
    if (fc == 1 && g == 8) return launchA_lc<1, 8>(lc, a, grid, st);
    if (fc == 2 && g == 4) return launchA_lc<2, 4>(lc, a, grid, st);
    if (fc == 4 && g == 2) return launchA_lc<4, 2>(lc, a, grid, st);
    return false;
}

// Phase-B launch kinds, each over its own item list: every regular-interior
// coord gathering (kB_All, b_stage 0), or waves staged in 128 / 256 LDS
// slots and the remaining regular waves (b_stage 1); then the other coords
// (kB_Gen).
enum { kB_All = 0, kB_S128 = 1, kB_S256 = 2, kB_Unl = 3, kB_Gen = 4, kBKinds = 5 };

struct PhaseB {
    BatchArgs a[kBKinds];
    int grid[kBKinds];
};

// persistent launches reset their claim counters first
template <class K>
void launchKind(const PhaseB &L, int kind, K kernel, hipStream_t st)
{
    const BatchArgs &a = L.a[kind];
    if (a.nitems == 0) return;
    if (a.persistent)
        hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, a.counter, kNumXcd);
    hipLaunchKernelGGL(kernel, dim3(L.grid[kind]), dim3(kBlock), 0, st, a);
}

template <int LC, int FC>
void launchB(const PhaseB &L, bool staged, hipStream_t st)
{
    if (staged) {
        launchKind(L, kB_S128, k_batch_patches_regular<LC, FC, kStagedOnly, 128>, st);
        // 256 slots one frame per chunk (LDS: 4 waves x FC x LC x 256 floats,
        // 28 KB at LC 7, so LDS allows as many blocks per CU as registers do)
        launchKind(L, kB_S256, k_batch_patches_regular<LC, 1, kStagedOnly, 256>, st);
        launchKind(L, kB_Unl, k_batch_patches_regular<LC, FC, kGatherUnlisted>, st);
    } else {
        launchKind(L, kB_All, k_batch_patches_regular<LC, FC, kGatherAll>, st);
    }
    // the few other coords (Gregory, boundary, bilinear): one frame per chunk
    // keeps the general kernel's register budget lower
    launchKind(L, kB_Gen, k_batch_patches<LC, 1>, st);
}

template <int FC>
bool launchB_lc(int lc, const PhaseB &L, bool staged, hipStream_t st)
{
    switch (lc) {
    case 1: launchB<1, FC>(L, staged, st); return true;
    case 2: launchB<2, FC>(L, staged, st); return true;
    case 3: launchB<3, FC>(L, staged, st); return true;
    case 4: launchB<4, FC>(L, staged, st); return true;
    case 5: launchB<5, FC>(L, staged, st); return true;
    case 6: launchB<6, FC>(L, staged, st); return true;
    case 7: launchB<7, FC>(L, staged, st); return true;
    case 8: launchB<8, FC>(L, staged, st); return true;
    default: return false;
    }
}

bool launchB_variant(int lc, int fc, bool staged, const PhaseB &L, hipStream_t st)
{
    if (fc == 1) return launchB_lc<1>(lc, L, staged, st);
    if (fc == 2) return launchB_lc<2>(lc, L, staged, st);
    if (fc == 4) return launchB_lc<4>(lc, L, staged, st);
    return false;
}

// Lists, per 64-coord wave of a patch segment, the distinct points of its
// regular-interior coords and each coord's 16 slot bytes (k_patch_slots
// count pass, scan, fill pass); the lists stay with the batch.
bool preparePatchStaging(osdhip_batch *b, BatchPatchSeg &g)
{
    const int nw = (g.ncoords + kWave - 1) / kWave;
    if (nw == 0) return true;
    int *nuniq = nullptr, *pts = nullptr, *ngen = nullptr;
    long long *listLen = nullptr, *ptsOff = nullptr;
    unsigned *slots = nullptr;
    void *tmp = nullptr;
    size_t tmpBytes = 0;
    const int grid = (nw + kWavesPerBlock - 1) / kWavesPerBlock;
    bool ok = hipMalloc(&nuniq, (size_t)nw * sizeof(int)) == hipSuccess &&
              hipMalloc(&ngen, (size_t)nw * sizeof(int)) == hipSuccess &&
              hipMalloc(&listLen, (size_t)nw * sizeof(long long)) == hipSuccess &&
              hipMalloc(&ptsOff, (size_t)nw * sizeof(long long)) == hipSuccess &&
              hipMalloc(&slots, (size_t)g.ncoords * 4 * sizeof(unsigned)) == hipSuccess;
    long long total = 0;
    if (ok) {
        hipLaunchKernelGGL(k_patch_slots<false>, dim3(grid), dim3(kBlock), 0, 0, g, nuniq, listLen,
                           ptsOff, pts, slots, ngen);
        ok = hipGetLastError() == hipSuccess &&
             hipcub::DeviceScan::ExclusiveSum(nullptr, tmpBytes, listLen, ptsOff, nw) == hipSuccess &&
             hipMalloc(&tmp, std::max<size_t>(tmpBytes, 1)) == hipSuccess &&
             hipcub::DeviceScan::ExclusiveSum(tmp, tmpBytes, listLen, ptsOff, nw) == hipSuccess;
    }
    long long lastOff = 0, lastLen = 0;
    std::vector<int> nu((size_t)nw), ng((size_t)nw);
    if (ok)
        ok = hipMemcpy(&lastOff, ptsOff + (nw - 1), sizeof(long long), hipMemcpyDeviceToHost) ==
                 hipSuccess &&
             hipMemcpy(&lastLen, listLen + (nw - 1), sizeof(long long), hipMemcpyDeviceToHost) ==
                 hipSuccess &&
             hipMemcpy(nu.data(), nuniq, (size_t)nw * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(ng.data(), ngen, (size_t)nw * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    total = lastOff + lastLen;
    if (ok && total > 0) ok = hipMalloc(&pts, (size_t)total * sizeof(int)) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_patch_slots<true>, dim3(grid), dim3(kBlock), 0, 0, g, nuniq, listLen,
                           ptsOff, pts, slots, ngen);
        ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    }
    if (tmp) (void)hipFree(tmp);
    if (listLen) (void)hipFree(listLen);
    for (void *p : { (void *)nuniq, (void *)ptsOff, (void *)pts, (void *)slots, (void *)ngen })
        if (p) b->prep.push_back(p);
    if (!ok) return false;
    for (int w = 0; w < nw; ++w)
        if (nu[w] >= 0) b->stagedCoords += std::min(kWave, g.ncoords - w * kWave);
    b->hNuniq.push_back(std::move(nu));
    b->hNgen.push_back(std::move(ng));
    g.nuniq = nuniq;
    g.ptsOff = ptsOff;
    g.pts = pts;
    g.slots = slots;
    g.ngen = ngen;
    return true;
}

// (Re)builds the phase-A item list for cfg.aItemSlices.
int buildItemsA(osdhip_batch *b)
{
    if (b->itemsA && b->itemsASlices == b->cfg.aItemSlices) return OSDHIP_OK;
    std::vector<BatchItem> items;
    // by stage; within a stage, segments with the most work first
    // (persistent balance)
    std::vector<int> order(b->nS);
    for (int i = 0; i < b->nS; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (b->sStage[x] != b->sStage[y]) return b->sStage[x] < b->sStage[y];
        return (long long)b->hS[x].frames * b->hS[x].nslices >
               (long long)b->hS[y].frames * b->hS[y].nslices;
    });
    const int per = b->cfg.aItemSlices;
    b->beginA.assign(b->nStages, 0);
    b->countA.assign(b->nStages, 0);
    for (int s : order) {
        const int st = b->sStage[s];
        if (b->countA[st] == 0) b->beginA[st] = (int)items.size();
        for (int f = 0; f < b->hS[s].nslices; f += per) {
            items.push_back({ s, f, std::min(per, b->hS[s].nslices - f), 0 });
            b->countA[st]++;
        }
    }
    if (b->itemsA) {
        (void)hipDeviceSynchronize();     // no queued kernel still reads the old list
        (void)hipFree(b->itemsA);
    }
    b->itemsA = nullptr;
    b->nItemsA = (int)items.size();
    b->itemsASlices = per;
    if (items.empty()) return OSDHIP_OK;
    if (hipMalloc(&b->itemsA, items.size() * sizeof(BatchItem)) != hipSuccess) return OSDHIP_ERR_ARGS;
    if (hipMemcpy(b->itemsA, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice) !=
        hipSuccess)
        return OSDHIP_ERR_ARGS;
    return OSDHIP_OK;
}

// Static: one block per item, rounded up to whole XCD groups (the block ->
// item map needs a multiple of 8).  Persistent: enough blocks to fill every
// CU a few times over.
int gridFor(const osdhip_batch *b, int nitems, bool persistent)
{
    if (!persistent) return grid_for_blocks(nitems);
    return grid_for_blocks(std::max(1, std::min(nitems, b->numCUs * 8)));
}

int enqueuePhaseA(osdhip_batch *b, int stage, hipStream_t st)
{
    const int n = b->countA[stage];
    if (n == 0) return OSDHIP_OK;
    BatchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.sseg = b->dS;
    a.items = b->itemsA + b->beginA[stage];
    a.nitems = n;
    a.persistent = b->cfg.aPersistent;
    a.counter = b->counters + kCounters * (2 * stage);
    if (!launchA_variant(b->tiled, b->length, b->cfg.aFrames, b->cfg.aGroup, a,
                         gridFor(b, n, a.persistent != 0), st))
        return OSDHIP_ERR_UNSUPPORTED;
    return last_error();
}

int enqueuePhaseB(osdhip_batch *b, int stage, hipStream_t st)
{
    PhaseB L;
    std::memset(&L, 0, sizeof(L));
    int total = 0;
    for (int k = 0; k < kBKinds; ++k) {
        BatchArgs &a = L.a[k];
        a.pseg = b->dP;
        a.items = b->itemsB + b->beginB[k][stage];
        a.nitems = b->countB[k][stage];
        a.persistent = b->cfg.bPersistent;
        a.counter = b->counters + kCounters * (2 * stage + 1);
        L.grid[k] = gridFor(b, a.nitems, a.persistent != 0);
        total += a.nitems;
    }
    if (total == 0) return OSDHIP_OK;
    if (!launchB_variant(b->length, b->cfg.bFrames, b->cfg.bStage != 0, L, st))
        return OSDHIP_ERR_UNSUPPORTED;
    return last_error();
}

// stage < 0: every stage in order
int enqueueAll(osdhip_batch *b, int phases, hipStream_t st, int stage = -1)
{
    int rc = OSDHIP_OK;
    if ((b->cfg.aPersistent && (phases & 1)) || (b->cfg.bPersistent && (phases & 2))) {
        const int nc = kCounters * 2 * b->nStages;
        hipLaunchKernelGGL(k_batch_reset, dim3((nc + 63) / 64), dim3(64), 0, st, b->counters, nc);
        rc = last_error();
    }
    for (int s = stage < 0 ? 0 : stage; rc == OSDHIP_OK && s < (stage < 0 ? b->nStages : stage + 1);
         ++s) {
        if (phases & 1) rc = enqueuePhaseA(b, s, st);
        if (rc == OSDHIP_OK && (phases & 2)) rc = enqueuePhaseB(b, s, st);
    }
    return rc;
}

void dropGraph(osdhip_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int capture(osdhip_batch *b)
{
    dropGraph(b);
    int rc = buildItemsA(b);
    if (rc != OSDHIP_OK) return rc;
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return last_error();
    rc = (int)hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (rc == OSDHIP_OK) {
        rc = enqueueAll(b, 3, cs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        if (rc == OSDHIP_OK) rc = (int)e;
        b->graph = g;
        if (rc == OSDHIP_OK) rc = (int)hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    }
    (void)hipStreamDestroy(cs);
    if (rc != OSDHIP_OK) dropGraph(b);
    return rc;
}

void freeBatch(osdhip_batch *b)
{
    if (!b) return;
    dropGraph(b);
    void *p[] = { b->dS, b->dP, b->itemsA, b->itemsB, b->counters };
    for (void *x : p)
        if (x) (void)hipFree(x);
    for (void *x : b->prep)
        if (x) (void)hipFree(x);
    delete b;
}

}  // namespace

extern "C" {

// osdhip_batch_create's body; *owned is the batch under construction, so
// that the caller frees it if a host allocation throws.
static osdhip_batch *batch_create(int length, int numStencilSegments,
                                  const osdhip_batch_stencils *stencils, int numPatchSegments,
                                  const osdhip_batch_patches *patches, void *stream,
                                  osdhip_batch **owned)
{
    if (length < 1 || length > 8 || numStencilSegments < 0 || numPatchSegments < 0 ||
        (numStencilSegments > 0 && !stencils) || (numPatchSegments > 0 && !patches))
        return nullptr;
    osdhip_batch *b = new (std::nothrow) osdhip_batch;
    if (!b) return nullptr;
    *owned = b;
    b->length = length;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        b->numCUs = std::max(1, prop.multiProcessorCount);

    for (int i = 0; i < numStencilSegments; ++i) {
        const osdhip_batch_stencils &s = stencils[i];
        const osdhip_stencil_table *t = s.table;
        if (!t || (t->layout != 0 && t->layout != 2) || t->nw != 1 || s.frames < 1 || !s.src ||
            !s.dst || s.srcVertexStride < length || s.dstVertexStride < length ||
            (i > 0 && (t->layout == 2) != b->tiled)) {
            freeBatch(b);
            return nullptr;
        }
        b->tiled = t->layout == 2;
        if (t->n == 0) continue;
        BatchStencilSeg g;
        std::memset(&g, 0, sizeof(g));
        g.rec = t->rec;
        g.sliceOff = t->sliceOff;
        g.sliceWidth = t->sliceWidth;
        g.rowSize = t->rowSize;
        g.perm = t->perm;
        g.nuniq = t->nuniq;
        g.ctrlOff = t->ctrlOff;
        g.ctrl = t->ctrl;
        g.src = s.src;
        g.srcV = s.srcVertexStride;
        g.srcF = s.srcFrameStride;
        g.dst = s.dst;
        g.dstV = s.dstVertexStride;
        g.dstF = s.dstFrameStride;
        g.nrows = t->n;
        g.nslices = t->nslices;
        g.frames = s.frames;
        if (s.stage < 0) {
            freeBatch(b);
            return nullptr;
        }
        b->hS.push_back(g);
        b->sStage.push_back(s.stage);
        b->nStages = std::max(b->nStages, s.stage + 1);
        b->tables.push_back(t);
        b->stencilRows += t->n;
        b->frameRows += (long long)t->n * s.frames;
        b->tableBytes += 4LL * t->n + 8LL * t->nnz;
    }
    std::vector<BatchItem> itemsB;
    for (int i = 0; i < numPatchSegments; ++i) {
        const osdhip_batch_patches &p = patches[i];
        if (p.numCoords < 0 || p.frames < 1 || !p.src || p.srcVertexStride < length ||
            (p.numCoords > 0 && (!p.coords || !p.arrays || !p.patchIndices || !p.params)) ||
            (p.P && p.outVertexStride < length) || (p.dPdu && p.outVertexStride < length) ||
            (p.dPdv && p.outVertexStride < length)) {
            freeBatch(b);
            return nullptr;
        }
        if (p.numCoords == 0) continue;
        if (p.stage < 0) {
            freeBatch(b);
            return nullptr;
        }
        BatchPatchSeg g;
        std::memset(&g, 0, sizeof(g));
        g.coords = p.coords;
        g.arrays = p.arrays;
        g.indices = p.patchIndices;
        g.params = p.params;
        g.src = p.src;
        g.srcV = p.srcVertexStride;
        g.srcF = p.srcFrameStride;
        g.out[0] = p.P;
        g.out[1] = p.dPdu;
        g.out[2] = p.dPdv;
        g.outV = p.outVertexStride;
        g.outF = p.outFrameStride;
        g.ncoords = p.numCoords;
        g.frames = p.frames;
        b->hP.push_back(g);
        b->pStage.push_back(p.stage);
        b->nStages = std::max(b->nStages, p.stage + 1);
        b->patchCoords += p.numCoords;
        b->patchFrames += (long long)p.numCoords * p.frames;
    }
    b->nS = (int)b->hS.size();
    b->nP = (int)b->hP.size();
    // per patch segment: point lists of its 64-coord waves (staged phase B)
    for (int i = 0; i < b->nP; ++i) {
        if (!preparePatchStaging(b, b->hP[i])) {
            freeBatch(b);
            return nullptr;
        }
    }
    // phase-B item lists per launch kind and stage, from the per-wave counts
    // of the staging pass (an item: kBlock coords = 4 waves)
    for (int k = 0; k < kBKinds; ++k) {
        b->beginB[k].assign(b->nStages, 0);
        b->countB[k].assign(b->nStages, 0);
        for (int st = 0; st < b->nStages; ++st) {
            b->beginB[k][st] = (int)itemsB.size();
            for (int seg = 0; seg < b->nP; ++seg) {
                if (b->pStage[seg] != st) continue;
                const int nc = b->hP[seg].ncoords;
                const std::vector<int> &nu = b->hNuniq[seg], &ng = b->hNgen[seg];
                for (int c = 0; c < nc; c += kBlock) {
                    bool take = k == kB_All;
                    for (int w = c / kWave; !take && w < (std::min(nc, c + kBlock) + kWave - 1) / kWave; ++w)
                        take = k == kB_S128 ? nu[w] > 0 && nu[w] <= 128
                             : k == kB_S256 ? nu[w] > 128
                             : k == kB_Unl ? nu[w] < 0
                             : ng[w] > 0;
                    if (take) itemsB.push_back({ seg, c, std::min(kBlock, nc - c), 0 });
                }
            }
            b->countB[k][st] = (int)itemsB.size() - b->beginB[k][st];
        }
    }
    b->nItemsB = (int)itemsB.size();
    b->hNuniq.clear();
    b->hNgen.clear();
    bool ok = hipMalloc(&b->counters, kCounters * 2 * (size_t)b->nStages * sizeof(unsigned)) ==
              hipSuccess;
    if (ok && b->nS)
        ok = hipMalloc(&b->dS, b->hS.size() * sizeof(BatchStencilSeg)) == hipSuccess &&
             hipMemcpy(b->dS, b->hS.data(), b->hS.size() * sizeof(BatchStencilSeg),
                       hipMemcpyHostToDevice) == hipSuccess;
    if (ok && b->nP)
        ok = hipMalloc(&b->dP, b->hP.size() * sizeof(BatchPatchSeg)) == hipSuccess &&
             hipMemcpy(b->dP, b->hP.data(), b->hP.size() * sizeof(BatchPatchSeg),
                       hipMemcpyHostToDevice) == hipSuccess;
    if (ok && b->nItemsB)
        ok = hipMalloc(&b->itemsB, itemsB.size() * sizeof(BatchItem)) == hipSuccess &&
             hipMemcpy(b->itemsB, itemsB.data(), itemsB.size() * sizeof(BatchItem),
                       hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = capture(b) == OSDHIP_OK;
    (void)stream;
    if (!ok) {
        freeBatch(b);
        return nullptr;
    }
    return b;
}

osdhip_batch *osdhip_batch_create(int length, int numStencilSegments,
                                  const osdhip_batch_stencils *stencils, int numPatchSegments,
                                  const osdhip_batch_patches *patches, void *stream)
{
    osdhip_batch *owned = nullptr;
    try {
        return batch_create(length, numStencilSegments, stencils, numPatchSegments, patches,
                            stream, &owned);
    } catch (...) {
        freeBatch(owned);
        return nullptr;
    }
}

void osdhip_batch_destroy(osdhip_batch *b) { freeBatch(b); }

int osdhip_batch_launch(osdhip_batch *b, void *stream)
{
    if (!b) return OSDHIP_ERR_ARGS;
    if (!b->exec) {
        int rc = capture(b);
        if (rc != OSDHIP_OK) return rc;
    }
    return (int)hipGraphLaunch(b->exec, (hipStream_t)stream);
}

int osdhip_batch_run(osdhip_batch *b, int phases, void *stream)
{
    if (!b || phases < 1 || phases > 3) return OSDHIP_ERR_ARGS;
    int rc = buildItemsA(b);
    if (rc != OSDHIP_OK) return rc;
    return enqueueAll(b, phases, (hipStream_t)stream);
}

int osdhip_batch_run_plain(osdhip_batch *b, void *stream)
{
    if (!b) return OSDHIP_ERR_ARGS;
    for (int s = 0; s < b->nStages; ++s) {
        const int rc = osdhip_batch_run_plain_stage(b, s, stream);
        if (rc != OSDHIP_OK) return rc;
    }
    return OSDHIP_OK;
}

int osdhip_batch_num_stages(const osdhip_batch *b) { return b ? b->nStages : OSDHIP_ERR_ARGS; }

int osdhip_batch_run_stage(osdhip_batch *b, int stage, int phases, void *stream)
{
    if (!b || stage < 0 || stage >= b->nStages || phases < 1 || phases > 3) return OSDHIP_ERR_ARGS;
    int rc = buildItemsA(b);
    if (rc != OSDHIP_OK) return rc;
    return enqueueAll(b, phases, (hipStream_t)stream, stage);
}

int osdhip_batch_run_plain_stage(osdhip_batch *b, int stage, void *stream)
{
    if (!b || stage < 0 || stage >= b->nStages) return OSDHIP_ERR_ARGS;
    const int L = b->length;
    // Osd buffer descriptors hold int strides; each frame is passed as its
    // own buffer (base pointer of the frame), as a caller with one buffer per
    // frame would
    const long long lim = 0x7fffffffLL;
    for (const BatchStencilSeg &g : b->hS)
        if (g.srcV > lim || g.dstV > lim) return OSDHIP_ERR_UNSUPPORTED;
    for (const BatchPatchSeg &g : b->hP)
        if (g.srcV > lim || g.outV > lim) return OSDHIP_ERR_UNSUPPORTED;
    for (int i = 0; i < b->nS; ++i) {
        const BatchStencilSeg &g = b->hS[i];
        if (b->sStage[i] != stage) continue;
        for (int f = 0; f < g.frames; ++f) {
            osdhip_buffer_desc sd = { 0, L, (int)g.srcV };
            osdhip_buffer_desc dd = { 0, L, (int)g.dstV };
            float *outs[1] = { g.dst + f * g.dstF };
            int rc = osdhip_stencil_table_eval(b->tables[i], g.src + f * g.srcF, sd, 1, outs, &dd,
                                               0, g.nrows, stream);
            if (rc != OSDHIP_OK) return rc;
        }
    }
    for (int i = 0; i < b->nP; ++i) {
        const BatchPatchSeg &g = b->hP[i];
        if (b->pStage[i] != stage) continue;
        for (int f = 0; f < g.frames; ++f) {
            osdhip_buffer_desc sd = { 0, L, (int)g.srcV };
            osdhip_buffer_desc od = { 0, L, (int)g.outV };
            osdhip_buffer_desc z = { 0, 0, 0 };
            float *o[3];
            for (int k = 0; k < 3; ++k) o[k] = g.out[k] ? g.out[k] + f * g.outF : nullptr;
            int rc = osdhip_eval_patches(g.src + f * g.srcF, sd, o[0], o[0] ? od : z, o[1],
                                         o[1] ? od : z, o[2], o[2] ? od : z, nullptr, z, nullptr,
                                         z, nullptr, z, g.ncoords, g.coords, g.arrays, g.indices,
                                         g.params, stream);
            if (rc != OSDHIP_OK) return rc;
        }
    }
    return OSDHIP_OK;
}

int osdhip_batch_set(osdhip_batch *b, const char *key, int value)
{
    if (!b || !key) return OSDHIP_ERR_ARGS;
    BatchConfig c = b->cfg;
    if (!std::strcmp(key, "a_frames")) c.aFrames = value;
    else if (!std::strcmp(key, "a_group")) c.aGroup = value;
    else if (!std::strcmp(key, "a_item_slices")) c.aItemSlices = value;
    else if (!std::strcmp(key, "a_persistent")) c.aPersistent = value;
    else if (!std::strcmp(key, "b_frames")) c.bFrames = value;
    else if (!std::strcmp(key, "b_persistent")) c.bPersistent = value;
    else if (!std::strcmp(key, "b_stage")) c.bStage = value;
    else return OSDHIP_ERR_UNSUPPORTED;
    // each knob within its own domain; the (frames, group) pair is checked
    // when the phase is launched
    if ((c.aFrames != 0 && c.aFrames != 1 && c.aFrames != 2 && c.aFrames != 4) ||
        (c.aGroup != 2 && c.aGroup != 4 && c.aGroup != 8) || c.aItemSlices < 1 ||
        c.aItemSlices > 1024 ||
        (c.aPersistent != 0 && c.aPersistent != 1) ||
        (c.bFrames != 1 && c.bFrames != 2 && c.bFrames != 4) ||
        (c.bPersistent != 0 && c.bPersistent != 1) || (c.bStage != 0 && c.bStage != 1))
        return OSDHIP_ERR_ARGS;
    b->cfg = c;
    dropGraph(b);          // re-captured on the next launch
    return OSDHIP_OK;
}

int osdhip_batch_get(const osdhip_batch *b, const char *key)
{
    if (!b || !key) return OSDHIP_ERR_ARGS;
    if (!std::strcmp(key, "a_frames")) return b->cfg.aFrames;
    if (!std::strcmp(key, "a_group")) return b->cfg.aGroup;
    if (!std::strcmp(key, "a_item_slices")) return b->cfg.aItemSlices;
    if (!std::strcmp(key, "a_persistent")) return b->cfg.aPersistent;
    if (!std::strcmp(key, "b_frames")) return b->cfg.bFrames;
    if (!std::strcmp(key, "b_persistent")) return b->cfg.bPersistent;
    if (!std::strcmp(key, "b_stage")) return b->cfg.bStage;
    if (!std::strcmp(key, "staged_coords"))
        return (int)std::min<long long>(b->stagedCoords, 0x7fffffffLL);
    return OSDHIP_ERR_UNSUPPORTED;
}

int osdhip_batch_info(const osdhip_batch *b, long long *stencilRows, long long *frameRows,
                      long long *patchCoords, long long *patchFrames, long long *tableBytes)
{
    if (!b) return OSDHIP_ERR_ARGS;
    if (stencilRows) *stencilRows = b->stencilRows;
    if (frameRows) *frameRows = b->frameRows;
    if (patchCoords) *patchCoords = b->patchCoords;
    if (patchFrames) *patchFrames = b->patchFrames;
    if (tableBytes) *tableBytes = b->tableBytes;
    return OSDHIP_OK;
}

int osdhip_batch_patch_staging(const osdhip_batch *b, int segment, int *nuniq, long long *ptsOff,
                               int *pts, unsigned *slots, long long *numPts)
{
    try {
        if (!b || segment < 0 || segment >= b->nP || !numPts) return OSDHIP_ERR_ARGS;
        const BatchPatchSeg &g = b->hP[(size_t)segment];
        const int nw = (g.ncoords + kWave - 1) / kWave;
        if (!g.nuniq || nw == 0) {
            *numPts = -1;
            return OSDHIP_OK;
        }
        std::vector<int> nu((size_t)nw);
        std::vector<long long> off((size_t)nw);
        hipError_t e = hipMemcpy(nu.data(), g.nuniq, (size_t)nw * sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(off.data(), g.ptsOff, (size_t)nw * sizeof(long long), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return (int)e;
        const long long total = off.back() + std::max(0, nu.back());
        *numPts = total;
        if (nuniq) std::copy(nu.begin(), nu.end(), nuniq);
        if (ptsOff) std::copy(off.begin(), off.end(), ptsOff);
        if (pts && total > 0 &&
            (e = hipMemcpy(pts, g.pts, (size_t)total * sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
            return (int)e;
        if (slots &&
            (e = hipMemcpy(slots, g.slots, (size_t)g.ncoords * 4 * sizeof(unsigned),
                           hipMemcpyDeviceToHost)) != hipSuccess)
            return (int)e;
        return OSDHIP_OK;
    } catch (...) {
        return OSDHIP_ERR_ARGS;
    }
}

int osdhip_batch_tune(osdhip_batch *b, int reps, void *stream, char *report, int reportBytes)
{
    try {
        if (!b || reps < 1) return OSDHIP_ERR_ARGS;
        hipStream_t st = (hipStream_t)stream;
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return last_error();
        std::string rep = "{";
        auto timePhase = [&](int phase, float &best) -> int {
            // one untimed run, then the median of reps timed runs
            int rc = osdhip_batch_run(b, phase, stream);
            std::vector<float> ms;
            for (int r = 0; rc == OSDHIP_OK && r < reps; ++r) {
                (void)hipEventRecord(e0, st);
                rc = osdhip_batch_run(b, phase, stream);
                (void)hipEventRecord(e1, st);
                if (rc == OSDHIP_OK) rc = (int)hipEventSynchronize(e1);
                float t = 0.0f;
                if (rc == OSDHIP_OK) rc = (int)hipEventElapsedTime(&t, e0, e1);
                ms.push_back(t);
            }
            if (rc != OSDHIP_OK) return rc;
            std::sort(ms.begin(), ms.end());
            best = ms[ms.size() / 2];
            return OSDHIP_OK;
        };
        int rc = OSDHIP_OK;
        char buf[160];
        // phase A: (frames per chunk, gather group) x item slices x scheduling
        if (b->nS) {
            const std::vector<std::pair<int, int>> fg =
                b->tiled ? std::vector<std::pair<int, int>>{ { 1, 4 }, { 1, 8 }, { 2, 4 }, { 2, 8 },
                                                             { 4, 4 }, { 4, 8 } }
                         : std::vector<std::pair<int, int>>{ { 0, 4 }, { 1, 4 }, { 1, 8 }, { 2, 4 },
                                                             { 4, 2 } };
            const int slices[3] = { 4, 16, 64 };
            BatchConfig bestc = b->cfg;
            float bestms = 1e30f;
            rep += "\"phase_a\":[";
            bool first = true;
            for (auto &v : fg)
                for (int sl : slices)
                    for (int pers = 0; pers < 2; ++pers) {
                        b->cfg.aFrames = v.first;
                        b->cfg.aGroup = v.second;
                        b->cfg.aItemSlices = sl;
                        b->cfg.aPersistent = pers;
                        float t = 0.0f;
                        rc = timePhase(1, t);
                        if (rc != OSDHIP_OK) break;
                        std::snprintf(buf, sizeof(buf),
                                      "%s{\"frames\":%d,\"group\":%d,\"slices\":%d,\"persistent\":%d,\"ms\":%.4f}",
                                      first ? "" : ",", v.first, v.second, sl, pers, t);
                        rep += buf;
                        first = false;
                        if (t < bestms) { bestms = t; bestc = b->cfg; }
                    }
            rep += "]";
            BatchConfig keep = b->cfg;
            keep.aFrames = bestc.aFrames;
            keep.aGroup = bestc.aGroup;
            keep.aItemSlices = bestc.aItemSlices;
            keep.aPersistent = bestc.aPersistent;
            b->cfg = keep;
        }
        // phase B: frames per chunk x scheduling (needs phase A's outputs: run it once first)
        if (rc == OSDHIP_OK && b->nP) {
            if (b->nS) rc = osdhip_batch_run(b, 1, stream);
            const int fcs[3] = { 1, 2, 4 };
            BatchConfig bestc = b->cfg;
            float bestms = 1e30f;
            rep += b->nS ? ",\"phase_b\":[" : "\"phase_b\":[";
            bool first = true;
            for (int fc : fcs)
                for (int pers = 0; rc == OSDHIP_OK && pers < 2; ++pers)
                    for (int stg = 0; rc == OSDHIP_OK && stg < 2; ++stg) {
                        b->cfg.bFrames = fc;
                        b->cfg.bPersistent = pers;
                        b->cfg.bStage = stg;
                        float t = 0.0f;
                        rc = timePhase(2, t);
                        if (rc != OSDHIP_OK) break;
                        std::snprintf(buf, sizeof(buf),
                                      "%s{\"frames\":%d,\"persistent\":%d,\"stage\":%d,\"ms\":%.4f}",
                                      first ? "" : ",", fc, pers, stg, t);
                        rep += buf;
                        first = false;
                        if (t < bestms) { bestms = t; bestc = b->cfg; }
                    }
            rep += "]";
            b->cfg.bFrames = bestc.bFrames;
            b->cfg.bPersistent = bestc.bPersistent;
            b->cfg.bStage = bestc.bStage;
        }
        rep += "}";
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc == OSDHIP_OK) rc = capture(b);
        if (report && reportBytes > 0) {
            std::strncpy(report, rep.c_str(), (size_t)reportBytes - 1);
            report[reportBytes - 1] = '\0';
        }
        return rc;
    } catch (const std::bad_alloc &) {
        return (int)hipErrorOutOfMemory;
    } catch (...) {
        return (int)hipErrorUnknown;
    }
}

}  // extern "C"

// ---- one-shot epochs (serving) ---------------------------------------------

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct EpochPlan {
    size_t segA = 0, segB = 0, itemsA = 0, itemsB = 0, total = 0;
    int nItemsA = 0, nItemsB = 0;
};

bool planEpoch(int length, int nS, const osdhip_batch_stencils *st, int nP,
               const osdhip_batch_patches *pt, int itemSlices, EpochPlan &p)
{
    if (length < 1 || length > 8 || nS < 0 || nP < 0 || itemSlices < 1 ||
        (nS && !st) || (nP && !pt))
        return false;
    long long ia = 0, ib = 0;
    for (int i = 0; i < nS; ++i) {
        const osdhip_stencil_table *t = st[i].table;
        if (!t || (t->layout != 0 && t->layout != 2) || t->nw != 1 || st[i].frames < 1 ||
            !st[i].src || !st[i].dst || st[i].srcVertexStride < length ||
            st[i].dstVertexStride < length || (t->layout == 2) != (st[0].table->layout == 2))
            return false;
        ia += (t->nslices + itemSlices - 1) / itemSlices;
    }
    for (int i = 0; i < nP; ++i) {
        const osdhip_batch_patches &q = pt[i];
        if (q.numCoords < 0 || q.frames < 1 || !q.src || q.srcVertexStride < length ||
            (q.numCoords > 0 && (!q.coords || !q.arrays || !q.patchIndices || !q.params)) ||
            ((q.P || q.dPdu || q.dPdv) && q.outVertexStride < length))
            return false;
        ib += (q.numCoords + kBlock - 1) / kBlock;
    }
    if (ia > 0x7fffffff || ib > 0x7fffffff) return false;
    p.nItemsA = (int)ia;
    p.nItemsB = (int)ib;
    p.segA = 0;
    p.segB = align256((size_t)nS * sizeof(BatchStencilSeg));
    p.itemsA = p.segB + align256((size_t)nP * sizeof(BatchPatchSeg));
    p.itemsB = p.itemsA + align256((size_t)ia * sizeof(BatchItem));
    p.total = p.itemsB + align256((size_t)ib * sizeof(BatchItem));
    return true;
}

}  // namespace

extern "C" {

long long osdhip_batch_epoch_bytes(int length, int numStencilSegments,
                                   const osdhip_batch_stencils *stencils, int numPatchSegments,
                                   const osdhip_batch_patches *patches, int itemSlices)
{
    EpochPlan p;
    if (!planEpoch(length, numStencilSegments, stencils, numPatchSegments, patches, itemSlices, p))
        return OSDHIP_ERR_ARGS;
    return (long long)p.total;
}

int osdhip_batch_epoch(int length, int numStencilSegments, const osdhip_batch_stencils *stencils,
                       int numPatchSegments, const osdhip_batch_patches *patches, int itemSlices,
                       void *staging, long long stagingBytes, void *scratch,
                       long long scratchBytes, void *stream)
{
    EpochPlan p;
    if (!planEpoch(length, numStencilSegments, stencils, numPatchSegments, patches, itemSlices, p))
        return OSDHIP_ERR_ARGS;
    if (!staging || !scratch || stagingBytes < (long long)p.total || scratchBytes < (long long)p.total)
        return OSDHIP_ERR_ARGS;
    char *h = (char *)staging;
    BatchStencilSeg *hs = (BatchStencilSeg *)(h + p.segA);
    BatchPatchSeg *hp = (BatchPatchSeg *)(h + p.segB);
    BatchItem *ha = (BatchItem *)(h + p.itemsA);
    BatchItem *hb = (BatchItem *)(h + p.itemsB);
    int na = 0, nb = 0;
    for (int i = 0; i < numStencilSegments; ++i) {
        const osdhip_batch_stencils &s = stencils[i];
        const osdhip_stencil_table *t = s.table;
        BatchStencilSeg g;
        std::memset(&g, 0, sizeof(g));
        g.rec = t->rec;
        g.sliceOff = t->sliceOff;
        g.sliceWidth = t->sliceWidth;
        g.rowSize = t->rowSize;
        g.perm = t->perm;
        g.nuniq = t->nuniq;
        g.ctrlOff = t->ctrlOff;
        g.ctrl = t->ctrl;
        g.src = s.src;
        g.srcV = s.srcVertexStride;
        g.srcF = s.srcFrameStride;
        g.dst = s.dst;
        g.dstV = s.dstVertexStride;
        g.dstF = s.dstFrameStride;
        g.nrows = t->n;
        g.nslices = t->nslices;
        g.frames = s.frames;
        hs[i] = g;
        for (int f = 0; f < t->nslices; f += itemSlices)
            ha[na++] = { i, f, std::min(itemSlices, t->nslices - f), 0 };
    }
    for (int i = 0; i < numPatchSegments; ++i) {
        const osdhip_batch_patches &q = patches[i];
        BatchPatchSeg g;
        std::memset(&g, 0, sizeof(g));
        g.coords = q.coords;
        g.arrays = q.arrays;
        g.indices = q.patchIndices;
        g.params = q.params;
        g.src = q.src;
        g.srcV = q.srcVertexStride;
        g.srcF = q.srcFrameStride;
        g.out[0] = q.P;
        g.out[1] = q.dPdu;
        g.out[2] = q.dPdv;
        g.outV = q.outVertexStride;
        g.outF = q.outFrameStride;
        g.ncoords = q.numCoords;
        g.frames = q.frames;
        hp[i] = g;
        for (int c = 0; c < q.numCoords; c += kBlock)
            hb[nb++] = { i, c, std::min(kBlock, q.numCoords - c), 0 };
    }
    hipStream_t st = (hipStream_t)stream;
    if (hipMemcpyAsync(scratch, staging, p.total, hipMemcpyHostToDevice, st) != hipSuccess)
        return last_error();
    char *d = (char *)scratch;
    BatchArgs a;
    std::memset(&a, 0, sizeof(a));
    if (na) {
        a.sseg = (const BatchStencilSeg *)(d + p.segA);
        a.items = (const BatchItem *)(d + p.itemsA);
        a.nitems = na;
        const bool tiled = stencils[0].table->layout == 2;
        if (!launchA_variant(tiled, length, 1, tiled ? 4 : 8, a, grid_for_blocks(na), st))
            return OSDHIP_ERR_UNSUPPORTED;
        int rc = last_error();
        if (rc != OSDHIP_OK) return rc;
    }
    if (nb) {
        // every item for the gathering regular launch and the general one
        PhaseB L;
        std::memset(&L, 0, sizeof(L));
        for (int k : { (int)kB_All, (int)kB_Gen }) {
            L.a[k].pseg = (const BatchPatchSeg *)(d + p.segB);
            L.a[k].items = (const BatchItem *)(d + p.itemsB);
            L.a[k].nitems = nb;
            L.grid[k] = grid_for_blocks(nb);
        }
        if (!launchB_variant(length, 1, false, L, st))
            return OSDHIP_ERR_UNSUPPORTED;
        int rc = last_error();
        if (rc != OSDHIP_OK) return rc;
    }
    return OSDHIP_OK;
}

}  // extern "C"
