// Batched evaluation: many meshes, each with many frames of control points,
// in one graph launch of two kernels.
//
//   phase A  k_batch_stencils  every stencil segment (one record-layout
//                              table applied to F frames): refinement
//                              tables and the patch-point tables that fill
//                              each mesh's vertex buffer [control | points];
//   phase B  k_batch_patches   every patch segment (one mesh's PatchCoords
//                              evaluated on its F frames), reading the
//                              vertex buffers phase A completed.
//
// Frames are extra right-hand-side columns: a wave loads a table slice's
// records once per chunk of FC frames (later chunks hit L1/L2), and a
// PatchCoord's basis is evaluated once for all frames.  Element (v, f, k)
// of a multi-frame buffer lives at base + v * vStride + f * fStride + k, so
// both frame-major and vertex-major (frames interleaved) layouts work.
//
// Work is cut into items of up to `itemSlices` 64-row slices (phase A) or
// kBlock coords (phase B).  Static scheduling maps one item to one block;
// persistent scheduling launches a fixed grid whose blocks claim items from
// a device counter (one vector atomic per item, by lane 0 of wave 0) until
// the list is exhausted -- every block exits when the counter passes the
// item count.
//
// Every sum accumulates w * x in the same order as the single-mesh kernels
// (k_stencils_rec / k_patches), so batched results equal one-mesh-per-launch
// results bit for bit.
#pragma once
#include "../../include/osdhip.h"
#include "common.h"
#include "patch_basis_device.h"
#include "stencil_sliced.h"
#include "stencil_tiled.h"

namespace osdhip {

struct BatchStencilSeg {
    const unsigned *rec;          // record layout, one weight array (R = 2)
    const long long *sliceOff;
    const int *sliceWidth;
    const int *rowSize;           // layout 0: padded to whole slices
    const int *perm;              // layout 2: slice row -> table row
    const int *nuniq;             // layout 2: listed control vertices per slice
    const long long *ctrlOff;
    const int *ctrl;
    const float *src;
    long long srcV, srcF;
    float *dst;
    long long dstV, dstF;
    int nrows, nslices, frames, pad;
};

struct BatchPatchSeg {
    const osdhip_patch_coord *coords;
    const osdhip_patch_array *arrays;
    const int *indices;
    const osdhip_patch_param *params;
    const float *src;
    long long srcV, srcF;
    float *out[3];                // P, dPdu, dPdv (NULL = skip)
    long long outV, outF;
    int ncoords, frames;
    // point staging of regular-interior coords, per wave of 64 coords
    // (NULL: gathers only): nuniq[w] distinct points listed at
    // pts[ptsOff[w] ...] (nuniq < 0: more than kPatchSlots), and per coord
    // the list positions of its 16 points, one byte each
    const int *nuniq;
    const long long *ptsOff;
    const int *pts;
    const unsigned *slots;        // 4 words per coord
    const int *ngen;              // per wave: coords that are not regular-interior
};

// Device views of the segments with global-qualified pointers (the
// descriptors are read from memory, so the compiler would otherwise treat
// their pointers as generic and emit flat accesses).
#define G_ OSDHIP_GLOBAL
struct GStencilSeg {
    const G_ unsigned *rec;
    const G_ long long *sliceOff;
    const G_ int *sliceWidth;
    const G_ int *rowSize;
    const G_ int *perm;
    const G_ int *nuniq;
    const G_ long long *ctrlOff;
    const G_ int *ctrl;
    const G_ float *src;
    long long srcV, srcF;
    G_ float *dst;
    long long dstV, dstF;
    int nrows, nslices, frames;
};

struct GPatchSeg {
    const G_ osdhip_patch_coord *coords;
    const G_ osdhip_patch_array *arrays;
    const G_ int *indices;
    const G_ osdhip_patch_param *params;
    const G_ float *src;
    long long srcV, srcF;
    G_ float *out[3];
    long long outV, outF;
    int ncoords, frames;
    const G_ int *nuniq;
    const G_ long long *ptsOff;
    const G_ int *pts;
    const G_ unsigned *slots;
    const G_ int *ngen;
};

#define C_ OSDHIP_CONST
__device__ __forceinline__ GStencilSeg gseg(const C_ BatchStencilSeg &s)
{
    GStencilSeg g;
    g.rec = gptr(s.rec);
    g.sliceOff = gptr(s.sliceOff);
    g.sliceWidth = gptr(s.sliceWidth);
    g.rowSize = gptr(s.rowSize);
    g.perm = gptr(s.perm);
    g.nuniq = gptr(s.nuniq);
    g.ctrlOff = gptr(s.ctrlOff);
    g.ctrl = gptr(s.ctrl);
    g.src = gptr(s.src);
    g.srcV = s.srcV;
    g.srcF = s.srcF;
    g.dst = gptr_w(s.dst);
    g.dstV = s.dstV;
    g.dstF = s.dstF;
    g.nrows = s.nrows;
    g.nslices = s.nslices;
    g.frames = s.frames;
    return g;
}

__device__ __forceinline__ GPatchSeg gseg(const C_ BatchPatchSeg &s)
{
    GPatchSeg g;
    g.coords = gptr(s.coords);
    g.arrays = gptr(s.arrays);
    g.indices = gptr(s.indices);
    g.params = gptr(s.params);
    g.src = gptr(s.src);
    g.srcV = s.srcV;
    g.srcF = s.srcF;
    for (int k = 0; k < 3; ++k) g.out[k] = s.out[k] ? gptr_w(s.out[k]) : nullptr;
    g.outV = s.outV;
    g.outF = s.outF;
    g.ncoords = s.ncoords;
    g.frames = s.frames;
    g.nuniq = gptr(s.nuniq);
    g.ptsOff = gptr(s.ptsOff);
    g.pts = gptr(s.pts);
    g.slots = gptr(s.slots);
    g.ngen = gptr(s.ngen);
    return g;
}

// Field-wise loads of the Osd structs through global pointers.
__device__ __forceinline__ osdhip_patch_coord ld(const G_ osdhip_patch_coord *p)
{
    osdhip_patch_coord c;
    c.arrayIndex = p->arrayIndex;
    c.patchIndex = p->patchIndex;
    c.vertIndex = p->vertIndex;
    c.s = p->s;
    c.t = p->t;
    return c;
}
__device__ __forceinline__ osdhip_patch_array ld(const G_ osdhip_patch_array *p)
{
    osdhip_patch_array a;
    a.regDesc = p->regDesc;
    a.desc = p->desc;
    a.numPatches = p->numPatches;
    a.indexBase = p->indexBase;
    a.stride = p->stride;
    a.primitiveIdBase = p->primitiveIdBase;
    return a;
}
__device__ __forceinline__ osdhip_patch_param ld(const G_ osdhip_patch_param *p)
{
    osdhip_patch_param q;
    q.field0 = p->field0;
    q.field1 = p->field1;
    q.sharpness = p->sharpness;
    return q;
}

// One (index, weight) record from a global pointer (non-temporal).
__device__ __forceinline__ void load_record_g(const OSDHIP_GLOBAL unsigned *p, int &idx, float &w)
{
    const v2i q = __builtin_nontemporal_load((const OSDHIP_GLOBAL v2i *)p);
    idx = q.x;
    w = __int_as_float(q.y);
}

struct BatchItem {
    int seg, first, count, pad;   // phase A: slices [first, first+count); B: coords
};

__device__ __forceinline__ BatchItem ld(const C_ BatchItem *p)
{
    BatchItem it;
    it.seg = p->seg;
    it.first = p->first;
    it.count = p->count;
    it.pad = 0;
    return it;
}

struct BatchArgs {
    const BatchStencilSeg *sseg;
    const BatchPatchSeg *pseg;
    const BatchItem *items;
    int nitems;
    int persistent;
    unsigned *counter;            // persistent: one claim counter per XCD queue
};

__global__ void k_batch_reset(unsigned *counters, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) counters[i] = 0u;
}

// Next item of this block.  Items are split into kNumXcd contiguous ranges,
// one per XCD (the hardware places block b on XCD b % 8), so neighbouring
// slices -- which share control points -- are processed by one XCD and
// share its L2.  Static: block b takes item xcd_logical_block() once.
// Persistent: block b claims items from its XCD's range through that
// range's counter and, when the range is exhausted, from the other ranges in
// turn; it returns -1 once every range is exhausted.  Must be reached by
// every thread of the block (it contains __syncthreads).  The result is
// wave-uniform.
__device__ __forceinline__ int batch_next_item(const BatchArgs &a, int iter)
{
    if (!a.persistent) {
        if (iter) return -1;
        const int lb = xcd_logical_block();
        return lb < a.nitems ? lb : -1;
    }
    __shared__ int slot, queue;
    __syncthreads();                          // previous item's readers are done
    if (threadIdx.x == 0) {
        if (iter == 0) queue = (int)(blockIdx.x % kNumXcd);
        const int per = (a.nitems + kNumXcd - 1) / kNumXcd;
        int it = -1;
        for (int k = 0; k < kNumXcd && it < 0; ++k) {
            const int q = (queue + k) % kNumXcd;
            const int lo = q * per, n = min(a.nitems, lo + per) - lo;
            if (n <= 0) continue;
            const int c = (int)atomicAdd(a.counter + q, 1u);
            if (c < n) {
                it = lo + c;
                queue = q;
            }
        }
        slot = it;
    }
    __syncthreads();
    // block-uniform: the item's descriptors are then read with scalar loads
    return __builtin_amdgcn_readfirstlane(slot);
}

template <int LC, int FC, int G>
__global__ __launch_bounds__(kBlock) void k_batch_stencils(BatchArgs a)
{
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wv = wave_in_block();
    for (int iter = 0;; ++iter) {
        const int itx = batch_next_item(a, iter);
        if (itx < 0) break;
        const BatchItem item = ld(cptr(a.items) + itx);
        const GStencilSeg sg = gseg(cptr(a.sseg)[item.seg]);
        for (int sl = wv; sl < item.count; sl += kWavesPerBlock) {
            const int s = item.first + sl;
            const int row = s * kWave + lane;
            const bool active = row < sg.nrows;
            const int sz = sg.rowSize[row];
            const int W = sg.sliceWidth[s];
            const G_ unsigned *rp = sg.rec + (size_t)(sg.sliceOff[s] + lane) * 2;
            for (int f0 = 0; f0 < sg.frames; f0 += FC) {
                float acc[FC][LC];
#pragma unroll
                for (int c = 0; c < FC; ++c)
#pragma unroll
                    for (int k = 0; k < LC; ++k) acc[c][k] = 0.0f;
                for (int j = 0; j < W; j += G) {
                    int ii[G];
                    float ww[1][G];
#pragma unroll
                    for (int u = 0; u < G; ++u)   // past the width: last column, not accumulated
                        load_record_g(rp + (size_t)min(j + u, W - 1) * (kWave * 2), ii[u], ww[0][u]);
                    float v[G][FC][LC];
#pragma unroll
                    for (int u = 0; u < G; ++u) {
                        const G_ float *p = sg.src + (size_t)(unsigned)ii[u] * sg.srcV +
                                                       (size_t)f0 * sg.srcF;
#pragma unroll
                        for (int c = 0; c < FC; ++c) {
                            if (f0 + c < sg.frames) {           // wave-uniform
#pragma unroll
                                for (int k = 0; k < LC; ++k) v[u][c][k] = p[c * sg.srcF + k];
                            } else {
#pragma unroll
                                for (int k = 0; k < LC; ++k) v[u][c][k] = 0.0f;
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < G; ++u) {
                        const bool take = j + u < sz;
#pragma unroll
                        for (int c = 0; c < FC; ++c)
#pragma unroll
                            for (int k = 0; k < LC; ++k) {
                                const float t = acc[c][k] + v[u][c][k] * ww[0][u];
                                acc[c][k] = take ? t : acc[c][k];
                            }
                    }
                }
                if (active) {
#pragma unroll
                    for (int c = 0; c < FC; ++c) {
                        if (f0 + c >= sg.frames) break;
                        G_ float *d =
                            sg.dst + (size_t)row * sg.dstV + (size_t)(f0 + c) * sg.dstF;
#pragma unroll
                        for (int k = 0; k < LC; ++k) d[k] = acc[c][k];
                    }
                }
            }
        }
    }
}

// A layout-2 slice with more distinct control vertices than LDS slots:
// records hold control indices (padding: kTilePad, skipped); every frame in
// turn gathers from memory and writes its results.
template <int LC, int G>
__device__ __forceinline__ void tiled_gather_slice(const GStencilSeg &sg, const G_ unsigned *rp, int W,
                                                int row, bool active)
{
    for (int f = 0; f < sg.frames; ++f) {
        float acc[LC];
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[k] = 0.0f;
        for (int j = 0; j < W; j += G) {
            int ii[G];
            float ww[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                load_record_g(rp + (size_t)min(j + u, W - 1) * (kWave * 2), ii[u], ww[u]);
                ii[u] = j + u < W ? ii[u] : (int)kTilePad;
            }
            float v[G][LC];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const unsigned ix = (unsigned)ii[u] == kTilePad ? 0u : (unsigned)ii[u];
                const G_ float *p = sg.src + (size_t)ix * sg.srcV + (size_t)f * sg.srcF;
#pragma unroll
                for (int k = 0; k < LC; ++k) v[u][k] = p[k];
            }
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const bool take = (unsigned)ii[u] != kTilePad;
#pragma unroll
                for (int k = 0; k < LC; ++k) {
                    const float t = acc[k] + v[u][k] * ww[u];
                    acc[k] = take ? t : acc[k];
                }
            }
        }
        if (active) {
            G_ float *d = sg.dst + (size_t)row * sg.dstV + (size_t)f * sg.dstF;
#pragma unroll
            for (int k = 0; k < LC; ++k) d[k] = acc[k];
        }
    }
}

// Records held in registers across every frame chunk (slices up to this
// width: the table then leaves memory once per step).
constexpr int kTiledRegRecords = 16;

// LDS floats per wave of the layout-2 batch kernel: the staging area
// (FC * LC values of 64 slots) doubles as the store transpose, whose rows
// have an odd stride so that both its writes and its reads are free of bank
// conflicts.
template <int LC, int FC>
struct TiledArea {
    static constexpr int Q = FC * LC;                 // values per row and chunk
    static constexpr int TP = Q | 1;                  // transpose row stride (odd)
    static constexpr int SEG = Q <= 8 ? 8 : Q <= 16 ? 16 : Q <= 32 ? 32 : 64;   // lanes per row
    static constexpr int RPI = kWave / SEG;           // rows per store instruction
    static constexpr int FLOATS = (Q > TP ? Q : TP) * kWave;
};

// Stores a chunk's results for vertex-major outputs (dstF == LC): each lane
// parks its row's FC * LC values in LDS, then every store instruction
// writes RPI rows, SEG consecutive lanes per row covering the row's
// contiguous run of FC * LC floats.
template <int LC, int FC>
__device__ __forceinline__ void tiled_store_runs(const GStencilSeg &sg, float *area,
                                                 const int *rowsL, const float (&acc)[FC][LC],
                                                 int f0)
{
    using A = TiledArea<LC, FC>;
    const int lane = (int)threadIdx.x & (kWave - 1);
    wave_lds_sync();                                   // compute reads of the area are done
#pragma unroll
    for (int c = 0; c < FC; ++c)
#pragma unroll
        for (int k = 0; k < LC; ++k) area[lane * A::TP + c * LC + k] = acc[c][k];
    wave_lds_sync();
    const int q = lane % A::SEG, hi = lane / A::SEG;
    const int nq = min(FC, sg.frames - f0) * LC;       // valid values per row
    G_ float *base = sg.dst + (size_t)f0 * LC + q;
#pragma unroll 4
    for (int it = 0; it < kWave / A::RPI; ++it) {
        const int r = it * A::RPI + hi;
        const int rr = rowsL[r];
        if (q < nq && rr >= 0) base[(size_t)rr * sg.dstV] = area[r * A::TP + q];
    }
}

// One listed layout-2 slice for every frame, FC frames per chunk: the
// wave stages its listed control vertices' frames into LDS (lane u ->
// slot u), then accumulates each row's entries from LDS.  RES: the slice's
// (up to N) records are loaded once into registers; otherwise they are read
// N per group for every chunk.
template <int LC, int FC, int N, bool RES, int G>
__device__ __forceinline__ void tiled_lds_slice(const GStencilSeg &sg, float *area, int *rowsL,
                                                const G_ unsigned *rp, int W, int U, int vtx,
                                                int row, bool active)
{
    const int lane = (int)threadIdx.x & (kWave - 1);
    float(*Ls)[kTileSlots] = (float(*)[kTileSlots])area;
    // vertex-major outputs with a row's frames adjacent: results leave
    // through LDS so that consecutive lanes store one row's run (measured
    // faster than per-lane stores for that layout; frame-major outputs keep
    // per-lane stores, which are faster still)
    const bool runs = sg.dstF == LC && sg.frames > 1;
    if (runs) rowsL[lane] = active ? row : -1;
    int slot[N];
    float ww[N];
    auto load = [&](int j0) {
#pragma unroll
        for (int u = 0; u < N; ++u) {
            // unconditional loads; past the width: the zero slot
            load_record_g(rp + (size_t)min(j0 + u, W - 1) * (kWave * 2), slot[u], ww[u]);
            slot[u] = j0 + u < W ? slot[u] : kTileZero;
            ww[u] = j0 + u < W ? ww[u] : 0.0f;
        }
    };
    if (RES) load(0);
    for (int f0 = 0; f0 < sg.frames; f0 += FC) {
        float acc[FC][LC];
#pragma unroll
        for (int c = 0; c < FC; ++c)
#pragma unroll
            for (int k = 0; k < LC; ++k) acc[c][k] = 0.0f;
        wave_lds_sync();                                 // previous chunk's reads are done
        if (lane < U) {
            const G_ float *p = sg.src + (size_t)(unsigned)vtx * sg.srcV + (size_t)f0 * sg.srcF;
#pragma unroll
            for (int c = 0; c < FC; ++c) {
                if (f0 + c < sg.frames) {                   // wave-uniform
#pragma unroll
                    for (int k = 0; k < LC; ++k) Ls[c * LC + k][lane] = p[c * sg.srcF + k];
                }
            }
        } else if (lane == kTileZero) {                  // the store transpose reuses the area
#pragma unroll
            for (int q = 0; q < FC * LC; ++q) Ls[q][kTileZero] = 0.0f;
        }
        wave_lds_sync();
        for (int j0 = 0; j0 < (RES ? 1 : W); j0 += N) {
            if (!RES) load(j0);
#pragma unroll
            for (int u = 0; u < N; ++u) {
                if (RES && u >= W) break;                   // wave-uniform
#pragma unroll
                for (int c = 0; c < FC; ++c)
#pragma unroll
                    for (int k = 0; k < LC; ++k)
                        acc[c][k] = acc[c][k] + Ls[c * LC + k][slot[u]] * ww[u];
                // one entry's LDS reads in flight at a time: hoisting all
                // of them would cut occupancy
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (runs) {
            tiled_store_runs<LC, FC>(sg, area, rowsL, acc, f0);
        } else if (active) {
#pragma unroll
            for (int c = 0; c < FC; ++c) {
                if (f0 + c >= sg.frames) break;
                G_ float *d = sg.dst + (size_t)row * sg.dstV + (size_t)(f0 + c) * sg.dstF;
#pragma unroll
                for (int k = 0; k < LC; ++k) d[k] = acc[c][k];
            }
        }
    }
}

// Phase A on locality-ordered tables (layout 2): per slice and chunk of FC
// frames, the wave stages the slice's listed control vertices (one per
// lane, FC frames each) into its LDS region, element (c, k) of slot u at
// lds[c * LC + k][u], then accumulates every entry from LDS; a slice's
// records stay in registers for all its chunks (slices of width up to
// kTiledRegRecords) or are re-read per chunk.  Results go to the table row
// perm[slice row].  Unlisted slices gather from memory.
template <int LC, int FC, int G>
__global__ __launch_bounds__(kBlock) void k_batch_stencils_tiled(BatchArgs a)
{
    __shared__ float lds[kWavesPerBlock][TiledArea<LC, FC>::FLOATS];
    __shared__ int rowsLds[kWavesPerBlock][kWave];
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wv = wave_in_block();
    for (int iter = 0;; ++iter) {
        const int itx = batch_next_item(a, iter);
        if (itx < 0) break;
        const BatchItem item = ld(cptr(a.items) + itx);
        const GStencilSeg sg = gseg(cptr(a.sseg)[item.seg]);
        for (int sl = wv; sl < item.count; sl += kWavesPerBlock) {
            const int s = item.first + sl;
            const int row = sg.perm[(long long)s * kWave + lane];
            const bool active = row < sg.nrows;
            const int W = sg.sliceWidth[s];
            const int U = sg.nuniq[s];
            const G_ unsigned *rp = sg.rec + (size_t)(sg.sliceOff[s] + lane) * 2;
            if (U < 0) {
                // rare: more distinct control vertices than LDS slots; one
                // frame at a time, gathers from memory
                tiled_gather_slice<LC, G>(sg, rp, W, row, active);
                continue;
            }
            const int vtx = lane < U ? sg.ctrl[sg.ctrlOff[s] + lane] : 0;
            if (W <= kTiledRegRecords)
                tiled_lds_slice<LC, FC, kTiledRegRecords, true, G>(sg, lds[wv], rowsLds[wv], rp, W,
                                                                   U, vtx, row, active);
            else
                tiled_lds_slice<LC, FC, G, false, G>(sg, lds[wv], rowsLds[wv], rp, W, U, vtx, row,
                                                     active);
        }
    }
}

// Phase A with the slice's records held in registers for all frames: a
// slice of width W <= 16 loads its (index, weight) records once, then every
// frame gathers and accumulates from registers, so the table leaves HBM
// exactly once per step whatever the frame count.  Wider slices take the
// chunked path (one frame per chunk, gather group 4).
template <int LC>
__global__ __launch_bounds__(kBlock) void k_batch_stencils_rr(BatchArgs a)
{
    constexpr int MAXW = 16, GG = 4;
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wv = wave_in_block();
    for (int iter = 0;; ++iter) {
        const int itx = batch_next_item(a, iter);
        if (itx < 0) break;
        const BatchItem item = ld(cptr(a.items) + itx);
        const GStencilSeg sg = gseg(cptr(a.sseg)[item.seg]);
        for (int sl = wv; sl < item.count; sl += kWavesPerBlock) {
            const int s = item.first + sl;
            const int row = s * kWave + lane;
            const bool active = row < sg.nrows;
            const int sz = sg.rowSize[row];
            const int W = sg.sliceWidth[s];
            const G_ unsigned *rp = sg.rec + (size_t)(sg.sliceOff[s] + lane) * 2;
            if (W <= MAXW) {
                int ii[MAXW];
                float ww[1][MAXW];
#pragma unroll
                for (int j = 0; j < MAXW; ++j)    // past the width: last column, not accumulated
                    load_record_g(rp + (size_t)min(j, W - 1) * (kWave * 2), ii[j], ww[0][j]);
                for (int f = 0; f < sg.frames; ++f) {
                    float acc[LC];
#pragma unroll
                    for (int k = 0; k < LC; ++k) acc[k] = 0.0f;
                    const G_ float *base = sg.src + (size_t)f * sg.srcF;
#pragma unroll
                    for (int j0 = 0; j0 < MAXW; j0 += GG) {
                        if (j0 < W) {                            // wave-uniform
                            float v[GG][LC];
#pragma unroll
                            for (int u = 0; u < GG; ++u) {
                                const G_ float *p = base + (size_t)(unsigned)ii[j0 + u] * sg.srcV;
#pragma unroll
                                for (int k = 0; k < LC; ++k) v[u][k] = p[k];
                            }
#pragma unroll
                            for (int u = 0; u < GG; ++u) {
                                const bool take = j0 + u < sz;
#pragma unroll
                                for (int k = 0; k < LC; ++k) {
                                    const float t = acc[k] + v[u][k] * ww[0][j0 + u];
                                    acc[k] = take ? t : acc[k];
                                }
                            }
                        }
                    }
                    if (active) {
                        G_ float *d = sg.dst + (size_t)row * sg.dstV + (size_t)f * sg.dstF;
#pragma unroll
                        for (int k = 0; k < LC; ++k) d[k] = acc[k];
                    }
                }
                continue;
            }
            for (int f = 0; f < sg.frames; ++f) {
                float acc[LC];
#pragma unroll
                for (int k = 0; k < LC; ++k) acc[k] = 0.0f;
                for (int j = 0; j < W; j += GG) {
                    int ii[GG];
                    float ww[1][GG];
#pragma unroll
                    for (int u = 0; u < GG; ++u)
                        load_record_g(rp + (size_t)min(j + u, W - 1) * (kWave * 2), ii[u], ww[0][u]);
                    float v[GG][LC];
#pragma unroll
                    for (int u = 0; u < GG; ++u) {
                        const G_ float *p = sg.src + (size_t)(unsigned)ii[u] * sg.srcV +
                                         (size_t)f * sg.srcF;
#pragma unroll
                        for (int k = 0; k < LC; ++k) v[u][k] = p[k];
                    }
#pragma unroll
                    for (int u = 0; u < GG; ++u) {
                        const bool take = j + u < sz;
#pragma unroll
                        for (int k = 0; k < LC; ++k) {
                            const float t = acc[k] + v[u][k] * ww[0][u];
                            acc[k] = take ? t : acc[k];
                        }
                    }
                }
                if (active) {
                    G_ float *d = sg.dst + (size_t)row * sg.dstV + (size_t)f * sg.dstF;
#pragma unroll
                    for (int k = 0; k < LC; ++k) d[k] = acc[k];
                }
            }
        }
    }
}

// Most points listed per wave of 64 coords (slot bytes: up to 256); the
// staged regular-patch kernel comes in 128- and 256-slot variants.
constexpr int kPatchSlots = 256;

// One wave per 64 coords of a patch segment (coords [64 w, 64 w + 64)):
// count pass (FILL false) -- distinct points of the wave's regular-interior
// coords (nuniq = -1 when more than kPatchSlots), list length and the
// number of other coords (ngen); fill pass -- the ascending point list and
// each regular-interior coord's 16 slot bytes.
// Distinct points are extracted in ascending order by repeated wave minima.
template <bool FILL>
__global__ __launch_bounds__(kBlock) void k_patch_slots(BatchPatchSeg sg, int *nuniq,
                                                        long long *listLen, const long long *ptsOff,
                                                        int *pts, unsigned *slots, int *ngen)
{
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int w = (int)blockIdx.x * kWavesPerBlock + wave_in_block();
    const int nw = (sg.ncoords + kWave - 1) / kWave;
    if (w >= nw) return;                                  // whole wave
    if (FILL && nuniq[w] < 0) return;
    const int i = w * kWave + lane;
    bool reg = false;
    int idx[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) idx[j] = 0;
    if (i < sg.ncoords) {
        const osdhip_patch_coord c = sg.coords[i];
        const osdhip_patch_array arr = sg.arrays[c.arrayIndex];
        const unsigned f1 = sg.params[c.patchIndex].field1;
        const int type = ((f1 >> 5) & 1u) ? arr.regDesc : arr.desc;
        reg = type == OSDHIP_PATCH_REGULAR && ((f1 >> 7) & 0x1fu) == 0;
        if (reg) {
            const int *cvs = sg.indices + arr.indexBase + arr.stride * (c.patchIndex - arr.primitiveIdBase);
#pragma unroll
            for (int j = 0; j < 16; ++j) idx[j] = cvs[j];
        }
    }
    unsigned sw[4] = { 0u, 0u, 0u, 0u };
    long long last = -1;
    int U = 0;
    bool over = false;
    for (;;) {
        unsigned m = 0xffffffffu;
        if (reg) {
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if ((long long)(unsigned)idx[j] > last) m = min(m, (unsigned)idx[j]);
        }
        m = wave_min_u(m);
        if (m == 0xffffffffu) break;
        if (U == kPatchSlots) {
            over = true;
            break;
        }
        if (FILL) {
            if (lane == 0) pts[ptsOff[w] + U] = (int)m;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (reg && (unsigned)idx[j] == m) sw[j / 4] |= (unsigned)U << (8 * (j % 4));
        }
        last = m;
        ++U;
    }
    if (!FILL) {
        const bool gen = i < sg.ncoords && !reg;
        const int ng = __popcll(__ballot(gen));
        if (lane == 0) {
            nuniq[w] = over ? -1 : U;
            listLen[w] = over ? 0 : U;
            ngen[w] = ng;
        }
        return;
    }
    if (reg) {
#pragma unroll
        for (int r = 0; r < 4; ++r) slots[(size_t)i * 4 + r] = sw[r];
    }
}

// acc[o][q][k] += x * w_o[j] for frames q of a chunk, reference order.
template <int LC, int FC>
__device__ __forceinline__ void batch_patch_point(const G_ float *p, long long fstride, int nq,
                                                  float wp, float wu, float wv,
                                                  float (&acc)[3][FC][LC])
{
#pragma unroll
    for (int q = 0; q < FC; ++q) {
        if (q < nq) {                                   // uniform per segment
#pragma unroll
            for (int k = 0; k < LC; ++k) {
                const float x = p[q * fstride + k];
                acc[0][q][k] += x * wp;
                acc[1][q][k] += x * wu;
                acc[2][q][k] += x * wv;
            }
        }
    }
}

// Phase B in two kernels over the same items, each with its own register
// budget:
//   k_batch_patches_regular  regular B-spline patches without boundary (nearly
//                            all coords): the 16 weights are formed in-line
//                            from the separable curves -- the same single
//                            products basis_bspline stores -- one row of 4
//                            points at a time, so few gathers are live;
//   k_batch_patches          every other coord (Gregory, boundary, bilinear),
//                            skipping the regular-interior ones.
__device__ __forceinline__ float pick4(const float (&v)[4], int r)
{
    return r == 0 ? v[0] : r == 1 ? v[1] : r == 2 ? v[2] : v[3];
}

// A regular-interior coord (REGULAR type, no boundary) and its parameters.
struct RegularCoord {
    bool ok;
    float s, t, k1;
    const G_ int *cvs;
};

__device__ __forceinline__ RegularCoord regular_coord(const GPatchSeg &sg, int i)
{
    RegularCoord rc;
    const osdhip_patch_coord c = ld(sg.coords + i);
    const osdhip_patch_array arr = ld(sg.arrays + c.arrayIndex);
    const osdhip_patch_param pp = ld(sg.params + c.patchIndex);
    const unsigned f1 = pp.field1;
    const int type = ((f1 >> 5) & 1u) ? arr.regDesc : arr.desc;
    const int boundary = (int)((f1 >> 7) & 0x1fu);
    rc.ok = type == OSDHIP_PATCH_REGULAR && boundary == 0;
    const int depth = (int)(f1 & 0xfu);
    const int nonquad = (int)((f1 >> 4) & 1u);
    const int pv = (int)((f1 >> 12) & 0x3ffu);
    const int pu = (int)((f1 >> 22) & 0x3ffu);
    const float frac_inv = (float)(1 << (depth - nonquad));
    rc.s = c.s * frac_inv - (float)pu;
    rc.t = c.t * frac_inv - (float)pv;
    rc.k1 = (float)(1 << depth);
    rc.cvs = sg.indices + arr.indexBase + arr.stride * (c.patchIndex - arr.primitiveIdBase);
    return rc;
}

template <int LC, int FC>
__device__ __forceinline__ void store_patch_chunk(const GPatchSeg &sg, int i, int f0, int nq,
                                                  float k1, const float (&acc)[3][FC][LC])
{
#pragma unroll
    for (int q = 0; q < FC; ++q) {
        if (q >= nq) break;
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            G_ float *d = sg.out[o];
            if (!d) continue;
            const float f = o == 0 ? 1.0f : k1;
            d += (size_t)i * sg.outV + (size_t)(f0 + q) * sg.outF;
#pragma unroll
            for (int k = 0; k < LC; ++k) d[k] = acc[o][q][k] * f;
        }
    }
}

// Waves of 64 coords in three modes, each launch with its own register
// budget: kGatherAll gathers every regular coord's points from memory;
// kStagedOnly takes the waves whose coords reference at most kPatchSlots
// distinct points (listed by k_patch_slots): the wave copies those points'
// frames into LDS -- element (q, k) of slot u at lds[q * LC + k][u], one or
// two slots per lane -- and reads each coord's 16 points from LDS through
// its slot bytes; kGatherUnlisted gathers for the other waves.  Same
// products and order in every mode.
enum { kGatherAll = 0, kStagedOnly = 1, kGatherUnlisted = 2 };

template <int LC, int FC, int MODE, int SLOTS = 128>
__global__ __launch_bounds__(kBlock) void k_batch_patches_regular(BatchArgs a)
{
    constexpr bool STAGED = MODE == kStagedOnly;
    __shared__ float lds[STAGED ? kWavesPerBlock : 1][STAGED ? FC * LC : 1][STAGED ? SLOTS : 1];
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wv = wave_in_block();
    for (int iter = 0;; ++iter) {
        const int itx = batch_next_item(a, iter);
        if (itx < 0) break;
        const BatchItem item = ld(cptr(a.items) + itx);
        const GPatchSeg sg = gseg(cptr(a.pseg)[item.seg]);
        const int i = item.first + (int)threadIdx.x;
        const bool inRange = i < item.first + item.count;
        // the wave's entry in the staging lists (items start at multiples of
        // kBlock coords); a wave past the segment's end has none
        const bool waveHasCoords = wv * kWave < item.count;
        const int U = (MODE != kGatherAll && sg.nuniq && waveHasCoords)
                          ? sg.nuniq[(item.first >> 6) + wv] : -1;
        // staged variants: 128 slots take U <= 128, 256 slots 128 < U <= 256
        if (MODE == kStagedOnly && (U < 0 || U > SLOTS || (SLOTS > 128 && U <= 128))) continue;
        if (MODE == kGatherUnlisted && U >= 0) continue;

        if (!STAGED) {                                    // gathers from memory
            if (!inRange) continue;
            const RegularCoord rc = regular_coord(sg, i);
            if (!rc.ok) continue;
            Curve cs, ct;
            bspline_curve(rc.s, cs);
            bspline_curve(rc.t, ct);
            for (int f0 = 0; f0 < sg.frames; f0 += FC) {
                const int nq = min(FC, sg.frames - f0);
                float acc[3][FC][LC];
#pragma unroll
                for (int o = 0; o < 3; ++o)
#pragma unroll
                    for (int q = 0; q < FC; ++q)
#pragma unroll
                        for (int k = 0; k < LC; ++k) acc[o][q][k] = 0.0f;
                const G_ float *base = sg.src + (size_t)f0 * sg.srcF;
#pragma unroll 1
                for (int r = 0; r < 4; ++r) {             // point j = 4 r + q
                    const float tb = pick4(ct.b, r), td = pick4(ct.d, r);
                    int idx[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) idx[q] = rc.cvs[4 * r + q];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        batch_patch_point<LC, FC>(base + (size_t)(unsigned)idx[q] * sg.srcV,
                                                  sg.srcF, nq, cs.b[q] * tb, cs.d[q] * tb,
                                                  cs.b[q] * td, acc);
                }
                store_patch_chunk<LC, FC>(sg, i, f0, nq, rc.k1, acc);
            }
            continue;
        }

        // staged wave: every lane takes part in the copies
        float(*Ls)[STAGED ? SLOTS : 1] = lds[STAGED ? wv : 0];
        RegularCoord rc;
        rc.ok = false;
        if (inRange) rc = regular_coord(sg, i);
        unsigned sw[4] = { 0u, 0u, 0u, 0u };
        if (rc.ok) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sw[r] = sg.slots[(size_t)i * 4 + r];
        }
        Curve cs, ct;
        bspline_curve(rc.ok ? rc.s : 0.0f, cs);
        bspline_curve(rc.ok ? rc.t : 0.0f, ct);
        const G_ int *lst = sg.pts + sg.ptsOff[(item.first >> 6) + wv];
        constexpr int PER = STAGED ? SLOTS / kWave : 1;   // slots per lane
        int vv[PER];
#pragma unroll
        for (int h = 0; h < PER; ++h) vv[h] = lane + h * kWave < U ? lst[lane + h * kWave] : 0;
        for (int f0 = 0; f0 < sg.frames; f0 += FC) {
            const int nq = min(FC, sg.frames - f0);
            wave_lds_sync();                              // previous chunk's reads are done
#pragma unroll
            for (int h = 0; h < PER; ++h) {
                const int u = lane + h * kWave;
                if (u < U) {
                    const G_ float *p = sg.src + (size_t)(unsigned)vv[h] * sg.srcV +
                                        (size_t)f0 * sg.srcF;
#pragma unroll
                    for (int q = 0; q < FC; ++q) {
                        if (q < nq) {                     // wave-uniform
#pragma unroll
                            for (int k = 0; k < LC; ++k) Ls[q * LC + k][u] = p[q * sg.srcF + k];
                        }
                    }
                }
            }
            wave_lds_sync();
            if (!rc.ok) continue;
            float acc[3][FC][LC];
#pragma unroll
            for (int o = 0; o < 3; ++o)
#pragma unroll
                for (int q = 0; q < FC; ++q)
#pragma unroll
                    for (int k = 0; k < LC; ++k) acc[o][q][k] = 0.0f;
            // the curves' values pass through an empty asm each chunk, so the
            // 48 weight products are formed per chunk instead of being held
            // in registers across the frame loop (they are the same products)
            float cb[4], cd[4], tbv[4], tdv[4];
            unsigned swv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                cb[r] = cs.b[r];
                cd[r] = cs.d[r];
                tbv[r] = ct.b[r];
                tdv[r] = ct.d[r];
                swv[r] = sw[r];
                asm volatile("" : "+v"(cb[r]), "+v"(cd[r]), "+v"(tbv[r]), "+v"(tdv[r]), "+v"(swv[r]));
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {                 // point j = 4 r + q, in order
                const float tb = tbv[r], td = tdv[r];
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) {
                    const int slot = (int)((swv[r] >> (8 * qq)) & 0xffu);
                    const float wp = cb[qq] * tb, wu = cd[qq] * tb, wvv = cb[qq] * td;
                    // every frame of the chunk without branches, so the point's
                    // LDS reads issue together (frames past the end are not stored)
                    float x[FC][LC];
#pragma unroll
                    for (int q = 0; q < FC; ++q)
#pragma unroll
                        for (int k = 0; k < LC; ++k) x[q][k] = Ls[q * LC + k][slot];
#pragma unroll
                    for (int q = 0; q < FC; ++q)
#pragma unroll
                        for (int k = 0; k < LC; ++k) {
                            acc[0][q][k] += x[q][k] * wp;
                            acc[1][q][k] += x[q][k] * wu;
                            acc[2][q][k] += x[q][k] * wvv;
                        }
                    // one point's reads in flight: the point's sums are taken
                    // as asm inputs before a memory clobber, so its products
                    // are formed before the next point's LDS reads issue
                    // (otherwise every point's 7 * FC values are held at once)
#pragma unroll
                    for (int o = 0; o < 3; ++o)
#pragma unroll
                        for (int q = 0; q < FC; ++q)
#pragma unroll
                            for (int k = 0; k < LC; ++k) asm volatile("" ::"v"(acc[o][q][k]));
                    asm volatile("" ::: "memory");
                }
            }
            store_patch_chunk<LC, FC>(sg, i, f0, nq, rc.k1, acc);
        }
    }
}

template <int LC, int FC>
__global__ __launch_bounds__(kBlock) void k_batch_patches(BatchArgs a)
{
    const int wv = wave_in_block();
    for (int iter = 0;; ++iter) {
        const int itx = batch_next_item(a, iter);
        if (itx < 0) break;
        const BatchItem item = ld(cptr(a.items) + itx);
        const GPatchSeg sg = gseg(cptr(a.pseg)[item.seg]);
        const int i = item.first + (int)threadIdx.x;
        if (i >= item.first + item.count) continue;
        // waves without such coords (counted when the batch was built)
        if (sg.ngen && sg.ngen[(item.first >> 6) + wv] == 0) continue;

        const osdhip_patch_coord c = ld(sg.coords + i);
        const osdhip_patch_array arr = ld(sg.arrays + c.arrayIndex);
        const osdhip_patch_param pp = ld(sg.params + c.patchIndex);
        const unsigned f1 = pp.field1;
        const int type = ((f1 >> 5) & 1u) ? arr.regDesc : arr.desc;
        const int boundary = (int)((f1 >> 7) & 0x1fu);
        if (type == OSDHIP_PATCH_REGULAR && boundary == 0) continue;   // the regular kernel's
        float s = c.s, t = c.t, sign;
        const float k1 = patch_normalize(type, f1, s, t, sign);
        const int n = patch_points(type);
        const G_ int *cvs = sg.indices + arr.indexBase + arr.stride * (c.patchIndex - arr.primitiveIdBase);
        Basis<true, false> B;
        if (n == 16) basis_bspline<true, false>(s, t, boundary, B);
        else if (n == 20) basis_gregory<true, false>(s, t, B);
        else if (n == 4) basis_linear<true, false>(s, t, B);
        else if (n == 12) basis_box_tri<true, false>(s, t, boundary, B);
        else if (n == 18) basis_gregory_tri<true, false>(s, t, B);
        else if (n == 3) basis_linear_tri<true, false>(s, t, B);
        // a negative (invalid) point index reads point 0, as in k_patches
        int cv[20];
#pragma unroll
        for (int j = 0; j < 20; ++j) cv[j] = j < n ? max(cvs[j], 0) : 0;

        for (int f0 = 0; f0 < sg.frames; f0 += FC) {
            const int nq = min(FC, sg.frames - f0);
            float acc[3][FC][LC];
#pragma unroll
            for (int o = 0; o < 3; ++o)
#pragma unroll
                for (int q = 0; q < FC; ++q)
#pragma unroll
                    for (int k = 0; k < LC; ++k) acc[o][q][k] = 0.0f;
            const G_ float *base = sg.src + (size_t)f0 * sg.srcF;
            // the point count follows from the type: 16 / 20 points without
            // per-point predicates, the others (at most 18) predicated, all
            // in the reference's point order
            if (n == 16 || n == 20) {
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    batch_patch_point<LC, FC>(base + (size_t)(unsigned)cv[j] * sg.srcV, sg.srcF,
                                              nq, B.p[j], B.ds[j], B.dt[j], acc);
                if (n == 20) {
#pragma unroll
                    for (int j = 16; j < 20; ++j)
                        batch_patch_point<LC, FC>(base + (size_t)(unsigned)cv[j] * sg.srcV,
                                                  sg.srcF, nq, B.p[j], B.ds[j], B.dt[j], acc);
                }
            } else {
#pragma unroll
                for (int j = 0; j < 18; ++j)
                    if (j < n)
                        batch_patch_point<LC, FC>(base + (size_t)(unsigned)cv[j] * sg.srcV,
                                                  sg.srcF, nq, B.p[j], B.ds[j], B.dt[j], acc);
            }
#pragma unroll
            for (int q = 0; q < FC; ++q) {
                if (q >= nq) break;
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    G_ float *d = sg.out[o];
                    if (!d) continue;
                    const float f = o == 0 ? 1.0f : k1;
                    d += (size_t)i * sg.outV + (size_t)(f0 + q) * sg.outF;
#pragma unroll
                    for (int k = 0; k < LC; ++k) d[k] = acc[o][q][k] * f;
                }
            }
        }
    }
}

}  // namespace osdhip
