// libosdhip: MI355X (gfx950, CDNA4) kernels for the Osd evaluator hot path
// and the C ABI declared in include/osdhip.h.
//
// Stencil evaluation is a sparse gather (CSR rows x a dense multi-column
// source), HBM-bound at < 0.3 FLOP/byte, so it uses no MFMA.  Two kernels:
//   * stencil_csr.h    - raw Far CSR tables (the raw EvalStencils entry
//                        points): per-wave LDS staging of the table slice;
//   * stencil_sliced.h - HipStencilTable's device layout (64-row column-major
//                        slices): coalesced streaming without LDS.
// All weight arrays of a derivative table (P, du, dv [, duu, duv, dvv])
// share the index stream and are applied in ONE launch.
//
// Patch evaluation (patch_kernels.h) computes the B-spline / Gregory /
// bilinear basis in registers per PatchCoord and gathers 16 / 20 / 4 points.
//
// Parity: built with -ffp-contract=off (no fused multiply-add) and every
// sum accumulates w * x in the reference's index order starting from 0, so
// stencil and patch results equal Osd::CpuEvaluator's bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/osdhip.h"
#include "common.h"
#include "patch_kernels.h"
#include "patch_map.h"
#include "tess.h"
#include "patch_sort.h"
#include "stencil_csr.h"
#include "stencil_sliced.h"
#include "stencil_tiled.h"
#include "table.h"
#include "table_build.h"

namespace osdhip {

template <int NWT, int G>
static void launch_rec_g(int lc, const RecArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_stencils_rec<NWT, 1, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_stencils_rec<NWT, 2, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_stencils_rec<NWT, 3, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_stencils_rec<NWT, 4, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_stencils_rec<NWT, 5, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_stencils_rec<NWT, 6, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_stencils_rec<NWT, 7, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: hipLaunchKernelGGL((k_stencils_rec<NWT, 8, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    }
}

template <int NWT>
static void launch_rec(int lc, int group, const RecArgs &a, int grid, hipStream_t st)
{
    if (group == 4) launch_rec_g<NWT, 4>(lc, a, grid, st);
    else if (group == 16) launch_rec_g<NWT, 16>(lc, a, grid, st);
    else launch_rec_g<NWT, 8>(lc, a, grid, st);
}

template <int G>
static void launch_rec_paired_g(int lc, const RecArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_stencils_rec_paired<1, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_stencils_rec_paired<2, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_stencils_rec_paired<3, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_stencils_rec_paired<4, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_stencils_rec_paired<5, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_stencils_rec_paired<6, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_stencils_rec_paired<7, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: hipLaunchKernelGGL((k_stencils_rec_paired<8, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    }
}

static void launch_rec_paired(int lc, int group, const RecArgs &a, int grid, hipStream_t st)
{
    if (group == 4) launch_rec_paired_g<4>(lc, a, grid, st);
    else if (group == 16) launch_rec_paired_g<16>(lc, a, grid, st);
    else launch_rec_paired_g<8>(lc, a, grid, st);
}

template <int NWT, int G>
static void launch_tiled_g(int lc, const TiledArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_stencils_tiled<NWT, 1, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_stencils_tiled<NWT, 2, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_stencils_tiled<NWT, 3, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_stencils_tiled<NWT, 4, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_stencils_tiled<NWT, 5, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_stencils_tiled<NWT, 6, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_stencils_tiled<NWT, 7, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: hipLaunchKernelGGL((k_stencils_tiled<NWT, 8, G>), dim3(grid), dim3(kBlock), 0, st, a); break;
    }
}

template <int NWT>
static void launch_tiled(int lc, int group, const TiledArgs &a, int grid, hipStream_t st)
{
    if (group == 4) launch_tiled_g<NWT, 4>(lc, a, grid, st);
    else launch_tiled_g<NWT, 8>(lc, a, grid, st);
}

template <int NW, bool VEC>
static void launch_csr(int lc, const StencilArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_stencils_csr<NW, 1, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_stencils_csr<NW, 2, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_stencils_csr<NW, 3, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_stencils_csr<NW, 4, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_stencils_csr<NW, 5, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_stencils_csr<NW, 6, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_stencils_csr<NW, 7, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: hipLaunchKernelGGL((k_stencils_csr<NW, 8, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    }
}

template <int NW>
static void launch_sliced(int lc, const SlicedArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_stencils_sliced<NW, 1>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_stencils_sliced<NW, 2>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_stencils_sliced<NW, 3>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_stencils_sliced<NW, 4>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_stencils_sliced<NW, 5>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_stencils_sliced<NW, 6>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_stencils_sliced<NW, 7>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: hipLaunchKernelGGL((k_stencils_sliced<NW, 8>), dim3(grid), dim3(kBlock), 0, st, a); break;
    }
}


static bool desc_ok(const osdhip_buffer_desc &d)
{
    return d.length > 0 && d.stride > 0 && d.offset >= 0 &&
           d.length <= d.stride - (d.offset % d.stride);
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

// Validates outputs against the source descriptor (CpuEvaluator's checks:
// osd/cpuEvaluator.cpp:46-47, 70-72); NULL outputs are skipped.
static bool outputs_ok(const osdhip_buffer_desc &sd, int nout, float *const *outs,
                       const osdhip_buffer_desc *ods, const float *const *ws)
{
    if (!desc_ok(sd)) return false;
    bool any = false;
    for (int w = 0; w < nout; ++w) {
        if (!outs[w]) continue;
        if (ods[w].length != sd.length || !desc_ok(ods[w]) || (ws && !ws[w])) return false;
        any = true;
    }
    return any;
}

// NW outputs over a raw CSR table.
template <int NW>
static int eval_csr(const float *src, osdhip_buffer_desc sd, float *const *outs,
                    const osdhip_buffer_desc *ods, const int *sizes, const int *offsets,
                    const int *indices, const float *const *ws, int start, int end, void *stream)
{
    if (end <= start) return OSDHIP_OK;
    if (!src || !sizes || !offsets || !indices || start < 0 || !ws[0]) return OSDHIP_ERR_ARGS;
    if (!outputs_ok(sd, NW, outs, ods, ws)) return OSDHIP_ERR_ARGS;
    const int count = end - start;
    StencilArgs a;
    std::memset(&a, 0, sizeof(a));
    a.srcStride = sd.stride;
    a.sizes = sizes;
    a.offsets = offsets;
    a.indices = indices;
    a.start = start;
    a.count = count;
    a.nblocks = (count + kBlock - 1) / kBlock;
    bool vec = aligned16(indices);
    for (int w = 0; w < NW; ++w) {
        // a skipped output still needs a readable weight stream for the
        // staging loop: reuse the P weights (their products are dropped)
        a.w[w] = (outs[w] && ws[w]) ? ws[w] : ws[0];
        vec = vec && aligned16(a.w[w]);
    }
    const int grid = grid_for_blocks(a.nblocks);
    hipStream_t st = (hipStream_t)stream;
    // Elements beyond 8 per vertex are handled by further launches over the
    // same stencils (windows of up to 8 floats).
    for (int e0 = 0; e0 < sd.length; e0 += 8) {
        const int lc = std::min(8, sd.length - e0);
        a.src = src + sd.offset + e0;
        for (int w = 0; w < NW; ++w) {
            a.dst[w] = outs[w] ? outs[w] + ods[w].offset + e0 : nullptr;
            a.dstStride[w] = outs[w] ? ods[w].stride : 0;
        }
        if (vec) launch_csr<NW, true>(lc, a, grid, st);
        else launch_csr<NW, false>(lc, a, grid, st);
        int rc = last_error();
        if (rc != OSDHIP_OK) return rc;
    }
    return OSDHIP_OK;
}

template <bool D1, bool D2>
static void launch_patches(int lc, const PatchArgs &a, hipStream_t st)
{
    const int grid = (a.n + kBlock - 1) / kBlock;
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_patches<D1, D2, 1>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_patches<D1, D2, 2>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_patches<D1, D2, 3>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_patches<D1, D2, 4>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_patches<D1, D2, 5>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_patches<D1, D2, 6>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_patches<D1, D2, 7>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: hipLaunchKernelGGL((k_patches<D1, D2, 8>), dim3(grid), dim3(kBlock), 0, st, a); break;
    }
}

}  // namespace osdhip

using namespace osdhip;


static void free_table(osdhip_stencil_table *t)
{
    if (!t) return;
    void *ptrs[] = { t->sizes, t->offsets, t->indices, t->w[0], t->w[1], t->w[2], t->w[3],
                     t->w[4], t->w[5], t->sliceOff, t->sliceWidth, t->rowSize, t->rec, t->sidx,
                     t->sw[0], t->sw[1], t->sw[2], t->sw[3], t->sw[4], t->sw[5], t->perm,
                     t->nuniq, t->ctrlOff, t->ctrl };
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete t;
}

template <typename T>
static bool dalloc(T **dptr, size_t count, long long &bytes)
{
    *dptr = nullptr;
    if (count == 0) return true;
    if (hipMalloc((void **)dptr, count * sizeof(T)) != hipSuccess) { *dptr = nullptr; return false; }
    bytes += (long long)(count * sizeof(T));
    return true;
}

template <typename T>
static bool upload(T **dptr, const T *host, size_t count, hipStream_t st, long long &bytes)
{
    *dptr = nullptr;
    if (count == 0) return true;
    if (hipMalloc((void **)dptr, count * sizeof(T)) != hipSuccess) { *dptr = nullptr; return false; }
    bytes += (long long)(count * sizeof(T));
    return hipMemcpyAsync(*dptr, host, count * sizeof(T), hipMemcpyHostToDevice, st) == hipSuccess;
}

// Exclusive prefix sum of n counts into out; returns the total (-1 on error).
template <typename T>
static long long exclusive_scan(const T *counts, T *out, int n, hipStream_t st)
{
    size_t bytes = 0;
    void *tmp = nullptr;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, counts, out, n, st) != hipSuccess ||
        hipMalloc(&tmp, std::max<size_t>(bytes, 1)) != hipSuccess)
        return -1;
    T last[2] = { 0, 0 };
    const bool ok =
        hipcub::DeviceScan::ExclusiveSum(tmp, bytes, counts, out, n, st) == hipSuccess &&
        hipMemcpyAsync(&last[0], out + (n - 1), sizeof(T), hipMemcpyDeviceToHost, st) == hipSuccess &&
        hipMemcpyAsync(&last[1], counts + (n - 1), sizeof(T), hipMemcpyDeviceToHost, st) ==
            hipSuccess &&
        hipStreamSynchronize(st) == hipSuccess;
    (void)hipFree(tmp);
    return ok ? (long long)last[0] + (long long)last[1] : -1;
}

// Layout 2 from the uploaded CSR arrays: dominant-vertex row order (stable
// radix sort), optionally regrouped by row size within windows of 128 or 512
// rows, per-slice widths and control lists (count pass, scans, fill pass).
static bool build_tiled(osdhip_stencil_table *t, hipStream_t st)
{
    const int n = t->n, S = t->nslices, rows = S * kWave;
    long long scratchBytes = 0;
    unsigned *keys = nullptr, *keysOut = nullptr;
    int *vals = nullptr;
    long long *entries = nullptr, *listLen = nullptr;
    void *sortTmp = nullptr;
    size_t sortBytes = 0;
    const int gr = (rows + kBlock - 1) / kBlock, gs = (S + kWavesPerBlock - 1) / kWavesPerBlock;
    const float *const *W = t->w;
    bool ok = dalloc(&keys, (size_t)rows, scratchBytes) && dalloc(&keysOut, (size_t)rows, scratchBytes) &&
              dalloc(&vals, (size_t)rows, scratchBytes) && dalloc(&entries, (size_t)S, scratchBytes) &&
              dalloc(&listLen, (size_t)S, scratchBytes) && dalloc(&t->perm, (size_t)rows, t->bytes) &&
              dalloc(&t->sliceWidth, (size_t)S, t->bytes) && dalloc(&t->sliceOff, (size_t)S, t->bytes) &&
              dalloc(&t->nuniq, (size_t)S, t->bytes) && dalloc(&t->ctrlOff, (size_t)S, t->bytes);
    if (ok) {
        hipLaunchKernelGGL(k_row_dominant, dim3(gr), dim3(kBlock), 0, st, t->sizes, t->offsets,
                           t->indices, W[0], n, rows, t->order == 1 ? 1 : 0, keys, vals);
        ok = hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, keys, keysOut, vals, t->perm, rows,
                                                0, 32, st) == hipSuccess &&
             hipMalloc(&sortTmp, std::max<size_t>(sortBytes, 1)) == hipSuccess &&
             hipcub::DeviceRadixSort::SortPairs(sortTmp, sortBytes, keys, keysOut, vals, t->perm,
                                                rows, 0, 32, st) == hipSuccess;
    }
#define OSDHIP_TILED(NWT, FILL, PERM)                                                           \
    hipLaunchKernelGGL((k_tiled_slices<NWT, FILL>), dim3(gs), dim3(kBlock), 0, st, PERM,        \
                       t->sizes, t->offsets, t->indices, W[0], W[1], W[2], W[3], W[4], W[5], n, \
                       S, t->sliceWidth, entries, t->nuniq, listLen, t->sliceOff, t->ctrlOff,   \
                       t->ctrl, t->rec)
    // count pass over a row order: slice widths and entries, control-list
    // counts; `listed` receives the slices that list their control vertices
    auto countPass = [&](const int *perm, int &listed) -> bool {
        if (t->nw == 1) OSDHIP_TILED(1, false, perm);
        else if (t->nw == 3) OSDHIP_TILED(3, false, perm);
        else OSDHIP_TILED(6, false, perm);
        std::unique_ptr<int[]> nu(new (std::nothrow) int[(size_t)S]);
        const bool good = hipGetLastError() == hipSuccess && nu &&
                          hipMemcpyAsync(nu.get(), t->nuniq, (size_t)S * sizeof(int),
                                         hipMemcpyDeviceToHost, st) == hipSuccess &&
                          hipStreamSynchronize(st) == hipSuccess;
        if (good)
            listed = (int)std::count_if(nu.get(), nu.get() + S, [](int u) { return u >= 0; });
        return good;
    };
    int listed = 0;
    if (ok) ok = countPass(t->perm, listed);
    // Size windows (order 0): of the dominant order
    // and its regroupings by size within 128- and 512-row windows, keep the
    // one whose slices hold the fewest record entries among those that list
    // as many slices as the dominant order.  The winner is kept in keysOut
    // (free after the sort), candidates are built in vals.
    t->sizeWindow = 0;
    if (ok && t->order == 0 && S > 1) {
        long long best = exclusive_scan(entries, t->sliceOff, S, st);
        ok = best >= 0;
        int *cand = vals, *keep = (int *)keysOut;
        const int base = listed;
        for (int K : { 128, 512 }) {
            if (!ok) break;
            const long long nwin = ((long long)rows + K - 1) / K;
            if (K == 128)
                hipLaunchKernelGGL(k_window_size_sort<128>, dim3((unsigned)nwin), dim3(kBlock), 0, st,
                                   t->perm, t->sizes, n, (long long)rows, cand);
            else
                hipLaunchKernelGGL(k_window_size_sort<512>, dim3((unsigned)nwin), dim3(kBlock), 0, st,
                                   t->perm, t->sizes, n, (long long)rows, cand);
            int l = 0;
            ok = hipGetLastError() == hipSuccess && countPass(cand, l);
            const long long e = ok ? exclusive_scan(entries, t->sliceOff, S, st) : -1;
            ok = ok && e >= 0;
            if (ok && l >= base && e < best) {
                best = e;
                t->sizeWindow = K;
                ok = hipMemcpyAsync(keep, cand, (size_t)rows * sizeof(int), hipMemcpyDeviceToDevice,
                                    st) == hipSuccess;
            }
        }
        if (ok && t->sizeWindow)
            ok = hipMemcpyAsync(t->perm, keep, (size_t)rows * sizeof(int), hipMemcpyDeviceToDevice,
                                st) == hipSuccess;
        if (ok) ok = countPass(t->perm, listed);      // the counts of the kept order
    }
    if (ok) t->listedSlices = listed;
    if (ok) {
        t->sliceEntries = exclusive_scan(entries, t->sliceOff, S, st);
        t->ctrlLen = exclusive_scan(listLen, t->ctrlOff, S, st);
        ok = t->sliceEntries >= 0 && t->ctrlLen >= 0 &&
             dalloc(&t->rec, (size_t)t->sliceEntries * rec_dwords(t->nw), t->bytes) &&
             dalloc(&t->ctrl, (size_t)t->ctrlLen, t->bytes);
    }
    if (ok) {
        if (t->nw == 1) OSDHIP_TILED(1, true, t->perm);
        else if (t->nw == 3) OSDHIP_TILED(3, true, t->perm);
        else OSDHIP_TILED(6, true, t->perm);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    }
#undef OSDHIP_TILED
    void *tmp[] = { keys, keysOut, vals, entries, listLen, sortTmp };
    for (void *p : tmp)
        if (p) (void)hipFree(p);
    return ok;
}

extern "C" {

int osdhip_eval_stencils(const float *src, osdhip_buffer_desc srcDesc,
                         float *dst, osdhip_buffer_desc dstDesc,
                         const int *sizes, const int *offsets,
                         const int *indices, const float *weights,
                         int start, int end, void *stream)
{
    if (end <= start) return OSDHIP_OK;
    if (srcDesc.length != dstDesc.length || !dst || !weights) return OSDHIP_ERR_ARGS;
    float *outs[1] = { dst };
    osdhip_buffer_desc ods[1] = { dstDesc };
    const float *ws[1] = { weights };
    return eval_csr<1>(src, srcDesc, outs, ods, sizes, offsets, indices, ws, start, end, stream);
}

int osdhip_eval_stencils_d1(const float *src, osdhip_buffer_desc srcDesc,
                            float *dst, osdhip_buffer_desc dstDesc,
                            float *du, osdhip_buffer_desc duDesc,
                            float *dv, osdhip_buffer_desc dvDesc,
                            const int *sizes, const int *offsets,
                            const int *indices, const float *weights,
                            const float *duWeights, const float *dvWeights,
                            int start, int end, void *stream)
{
    if (end <= start) return OSDHIP_OK;
    float *outs[3] = { dst, du, dv };
    osdhip_buffer_desc ods[3] = { dstDesc, duDesc, dvDesc };
    const float *ws[3] = { weights, duWeights, dvWeights };
    return eval_csr<3>(src, srcDesc, outs, ods, sizes, offsets, indices, ws, start, end, stream);
}

int osdhip_eval_stencils_d2(const float *src, osdhip_buffer_desc srcDesc,
                            float *dst, osdhip_buffer_desc dstDesc,
                            float *du, osdhip_buffer_desc duDesc,
                            float *dv, osdhip_buffer_desc dvDesc,
                            float *duu, osdhip_buffer_desc duuDesc,
                            float *duv, osdhip_buffer_desc duvDesc,
                            float *dvv, osdhip_buffer_desc dvvDesc,
                            const int *sizes, const int *offsets,
                            const int *indices, const float *weights,
                            const float *duWeights, const float *dvWeights,
                            const float *duuWeights, const float *duvWeights,
                            const float *dvvWeights,
                            int start, int end, void *stream)
{
    if (end <= start) return OSDHIP_OK;
    float *outs[6] = { dst, du, dv, duu, duv, dvv };
    osdhip_buffer_desc ods[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
    const float *ws[6] = { weights, duWeights, dvWeights, duuWeights, duvWeights, dvvWeights };
    return eval_csr<6>(src, srcDesc, outs, ods, sizes, offsets, indices, ws, start, end, stream);
}

void osdhip_table_options_init(osdhip_table_options *o)
{
    if (!o) return;
    o->layout = 0;
    o->keepCsr = 1;
    o->group = 8;
    o->kernel = 0;
    o->order = 0;
}

static bool table_options_ok(const osdhip_table_options &o)
{
    return o.layout >= 0 && o.layout <= 3 && (o.keepCsr == 0 || o.keepCsr == 1) &&
           (o.group == 4 || o.group == 8 || o.group == 16) && (o.kernel == 0 || o.kernel == 1) &&
           !(o.kernel == 1 && !o.keepCsr) && (o.order >= 0 && o.order <= 2);
}

osdhip_stencil_table *osdhip_stencil_table_create(int numStencils, const int *sizes,
                                                  const int *offsets, const int *indices,
                                                  int numWeightArrays,
                                                  const float *const *weights, void *stream)
{
    return osdhip_stencil_table_create_ex(numStencils, sizes, offsets, indices, numWeightArrays,
                                          weights, nullptr, stream);
}

osdhip_stencil_table *osdhip_stencil_table_create_ex(int numStencils, const int *sizes,
                                                     const int *offsets, const int *indices,
                                                     int numWeightArrays,
                                                     const float *const *weights,
                                                     const osdhip_table_options *options,
                                                     void *stream)
{
    osdhip_table_options opt;
    osdhip_table_options_init(&opt);
    if (options) opt = *options;
    if (!table_options_ok(opt)) return nullptr;
    if (numStencils < 0 || (numStencils > 0 && (!sizes || !indices || !weights)) ||
        (numWeightArrays != 1 && numWeightArrays != 3 && numWeightArrays != 6))
        return nullptr;
    if (opt.layout == 3 && numWeightArrays != 1) return nullptr;   // pairs of 8-byte records
    for (int k = 0; k < numWeightArrays; ++k)
        if (numStencils > 0 && !weights[k]) return nullptr;
    osdhip_stencil_table *t = new (std::nothrow) osdhip_stencil_table;
    if (!t) return nullptr;
    t->n = numStencils;
    t->nw = numWeightArrays;
    t->keepCsr = opt.keepCsr;
    t->group = opt.group;
    t->kernel = opt.kernel;
    t->order = opt.order;
    hipStream_t st = (hipStream_t)stream;

    // CSR mirror (offsets regenerated when absent: exclusive prefix of sizes)
    std::unique_ptr<int[]> offBuf(new (std::nothrow) int[(size_t)std::max(numStencils, 1)]);
    if (!offBuf) { free_table(t); return nullptr; }
    int *const off = offBuf.get();
    long long run = 0;
    for (int i = 0; i < numStencils; ++i) {
        if (sizes[i] < 0) { free_table(t); return nullptr; }
        off[i] = offsets ? offsets[i] : (int)run;
        run += sizes[i];
    }
    long long nnz = 0;
    for (int i = 0; i < numStencils; ++i) nnz = std::max(nnz, (long long)off[i] + sizes[i]);
    t->nnz = nnz;
    bool ok = upload(&t->sizes, sizes, (size_t)numStencils, st, t->bytes) &&
              upload(&t->offsets, off, (size_t)numStencils, st, t->bytes) &&
              upload(&t->indices, indices, (size_t)nnz, st, t->bytes);
    for (int k = 0; ok && k < numWeightArrays; ++k)
        ok = upload(&t->w[k], weights[k], (size_t)nnz, st, t->bytes);

    // slice layout, built on the device from the CSR mirror
    const int S = (numStencils + kWave - 1) / kWave;
    const int rows = S * kWave;
    t->nslices = S;
    t->layout = opt.layout;
    long long *entries = nullptr;
    void *scanTmp = nullptr;
    size_t scanBytes = 0;
    long long dummy = 0;
    if (ok && S > 0 && t->layout == 2) {
        ok = build_tiled(t, st);
    } else if (ok && S > 0) {
        ok = dalloc(&t->sliceOff, (size_t)S, t->bytes) && dalloc(&t->sliceWidth, (size_t)S, t->bytes) &&
             dalloc(&t->rowSize, (size_t)rows, t->bytes) && dalloc(&entries, (size_t)S, dummy);
        if (ok) {
            const int gs = (S + kBlock - 1) / kBlock, gr = (rows + kBlock - 1) / kBlock;
            hipLaunchKernelGGL(k_slice_widths, dim3(gs), dim3(kBlock), 0, st, t->sizes, numStencils,
                               S, t->sliceWidth, entries, t->layout == 3 ? 1 : 0);
            hipLaunchKernelGGL(k_pad_sizes, dim3(gr), dim3(kBlock), 0, st, t->sizes, numStencils,
                               rows, t->rowSize);
            ok = hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, entries, t->sliceOff, S, st) ==
                     hipSuccess &&
                 hipMalloc(&scanTmp, scanBytes) == hipSuccess &&
                 hipcub::DeviceScan::ExclusiveSum(scanTmp, scanBytes, entries, t->sliceOff, S, st) ==
                     hipSuccess;
        }
        long long lastOff = 0, lastEntries = 0;
        if (ok)
            ok = hipMemcpyAsync(&lastOff, t->sliceOff + (S - 1), sizeof(long long),
                                hipMemcpyDeviceToHost, st) == hipSuccess &&
                 hipMemcpyAsync(&lastEntries, entries + (S - 1), sizeof(long long),
                                hipMemcpyDeviceToHost, st) == hipSuccess &&
                 hipStreamSynchronize(st) == hipSuccess;
        t->sliceEntries = lastOff + lastEntries;
        const int gr = (rows + kBlock - 1) / kBlock;
        const float *const *W = t->w;
        if (ok && t->layout == 3) {
            ok = dalloc(&t->rec, (size_t)t->sliceEntries * 2, t->bytes);
            if (ok)
                hipLaunchKernelGGL(k_fill_records_paired, dim3(gr), dim3(kBlock), 0, st, t->sizes,
                                   t->offsets, t->indices, W[0], numStencils, rows, t->sliceOff,
                                   t->sliceWidth, t->rec);
        } else if (ok && t->layout == 0) {
            ok = dalloc(&t->rec, (size_t)t->sliceEntries * rec_dwords(numWeightArrays), t->bytes);
            if (ok) {
                if (numWeightArrays == 1)
                    hipLaunchKernelGGL(k_fill_records<1>, dim3(gr), dim3(kBlock), 0, st, t->sizes,
                                       t->offsets, t->indices, W[0], W[1], W[2], W[3], W[4], W[5],
                                       numStencils, rows, t->sliceOff, t->sliceWidth, t->rec);
                else if (numWeightArrays == 3)
                    hipLaunchKernelGGL(k_fill_records<3>, dim3(gr), dim3(kBlock), 0, st, t->sizes,
                                       t->offsets, t->indices, W[0], W[1], W[2], W[3], W[4], W[5],
                                       numStencils, rows, t->sliceOff, t->sliceWidth, t->rec);
                else
                    hipLaunchKernelGGL(k_fill_records<6>, dim3(gr), dim3(kBlock), 0, st, t->sizes,
                                       t->offsets, t->indices, W[0], W[1], W[2], W[3], W[4], W[5],
                                       numStencils, rows, t->sliceOff, t->sliceWidth, t->rec);
            }
        } else if (ok) {
            ok = dalloc(&t->sidx, (size_t)t->sliceEntries, t->bytes);
            for (int k = 0; ok && k < numWeightArrays; ++k)
                ok = dalloc(&t->sw[k], (size_t)t->sliceEntries, t->bytes);
            if (ok)
                hipLaunchKernelGGL(k_fill_split, dim3(gr), dim3(kBlock), 0, st, t->sizes, t->offsets,
                                   t->indices, W[0], W[1], W[2], W[3], W[4], W[5], numWeightArrays,
                                   numStencils, rows, t->sliceOff, t->sliceWidth, t->sidx, t->sw[0],
                                   t->sw[1], t->sw[2], t->sw[3], t->sw[4], t->sw[5]);
        }
        ok = ok && hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    }
    if (entries) (void)hipFree(entries);
    if (scanTmp) (void)hipFree(scanTmp);
    if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
    if (!ok) { free_table(t); return nullptr; }
    if (!t->keepCsr) {
        // slice layout only: the CSR arrays were needed to build it
        void *csr[] = { t->offsets, t->indices, t->w[0], t->w[1], t->w[2], t->w[3], t->w[4], t->w[5] };
        for (void *p : csr)
            if (p) (void)hipFree(p);
        t->bytes -= 4LL * numStencils + 4LL * nnz * (1 + numWeightArrays);
        t->offsets = t->indices = nullptr;
        for (int k = 0; k < 6; ++k) t->w[k] = nullptr;
    }
    return t;
}

void osdhip_stencil_table_destroy(osdhip_stencil_table *t) { free_table(t); }

int osdhip_stencil_table_set(osdhip_stencil_table *t, const char *key, int value)
{
    if (!t || !key) return OSDHIP_ERR_ARGS;
    osdhip_table_options o;
    if (osdhip_stencil_table_options(t, &o) != OSDHIP_OK) return OSDHIP_ERR_ARGS;
    int *field = nullptr;
    if (!std::strcmp(key, "group")) { o.group = value; field = &t->group; }
    else if (!std::strcmp(key, "kernel")) { o.kernel = value; field = &t->kernel; }
    else return OSDHIP_ERR_UNSUPPORTED;
    if (!table_options_ok(o) || (o.kernel == 1 && !t->indices)) return OSDHIP_ERR_ARGS;
    const int old = *field;
    *field = value;
    return old;
}

int osdhip_stencil_table_options(const osdhip_stencil_table *t, osdhip_table_options *o)
{
    if (!t || !o) return OSDHIP_ERR_ARGS;
    o->layout = t->layout;
    o->keepCsr = t->keepCsr;
    o->group = t->group;
    o->kernel = t->kernel;
    o->order = t->order;
    return OSDHIP_OK;
}

int osdhip_stencil_table_info(const osdhip_stencil_table *t, int *numStencils,
                              long long *numEntries, int *numWeightArrays, long long *deviceBytes)
{
    if (!t) return OSDHIP_ERR_ARGS;
    if (numStencils) *numStencils = t->n;
    if (numEntries) *numEntries = t->nnz;
    if (numWeightArrays) *numWeightArrays = t->nw;
    if (deviceBytes) *deviceBytes = t->bytes;
    return OSDHIP_OK;
}

int osdhip_stencil_table_layout(const osdhip_stencil_table *t, int *layout, int *numSlices,
                                int *listedSlices, long long *listEntries)
{
    if (!t) return OSDHIP_ERR_ARGS;
    if (layout) *layout = t->layout;
    if (numSlices) *numSlices = t->nslices;
    if (listedSlices) *listedSlices = t->listedSlices;
    if (listEntries) *listEntries = t->ctrlLen;
    return OSDHIP_OK;
}

int osdhip_stencil_table_slices(const osdhip_stencil_table *t, long long *recordEntries,
                                int *sizeWindow)
{
    if (!t) return OSDHIP_ERR_ARGS;
    if (recordEntries) *recordEntries = t->sliceEntries;
    if (sizeWindow) *sizeWindow = t->sizeWindow;
    return OSDHIP_OK;
}

int osdhip_stencil_table_csr(const osdhip_stencil_table *t, const int **sizes, const int **offsets,
                             const int **indices, const float **weights)
{
    if (!t) return OSDHIP_ERR_ARGS;
    if (sizes) *sizes = t->sizes;
    if (offsets) *offsets = t->offsets;
    if (indices) *indices = t->indices;
    if (weights)
        for (int k = 0; k < 6; ++k) weights[k] = t->w[k];
    return OSDHIP_OK;
}

int osdhip_stencil_table_eval(const osdhip_stencil_table *t, const float *src,
                              osdhip_buffer_desc srcDesc, int numOutputs, float *const *outputs,
                              const osdhip_buffer_desc *outputDescs, int start, int end,
                              void *stream)
{
    if (!t || !outputs || !outputDescs) return OSDHIP_ERR_ARGS;
    if (numOutputs != 1 && numOutputs != 3 && numOutputs != 6) return OSDHIP_ERR_ARGS;
    if (numOutputs > t->nw) return OSDHIP_ERR_ARGS;
    end = std::min(end, t->n);
    if (end <= start) return OSDHIP_OK;
    if (!src || start < 0 || !outputs_ok(srcDesc, numOutputs, outputs, outputDescs, nullptr))
        return OSDHIP_ERR_ARGS;
    if (t->kernel == 1 && t->indices) {
        const float *ws[6] = { t->w[0], t->w[1], t->w[2], t->w[3], t->w[4], t->w[5] };
        if (numOutputs == 1)
            return eval_csr<1>(src, srcDesc, outputs, outputDescs, t->sizes, t->offsets, t->indices,
                               ws, start, end, stream);
        if (numOutputs == 3)
            return eval_csr<3>(src, srcDesc, outputs, outputDescs, t->sizes, t->offsets, t->indices,
                               ws, start, end, stream);
        return eval_csr<6>(src, srcDesc, outputs, outputDescs, t->sizes, t->offsets, t->indices,
                           ws, start, end, stream);
    }
    const int sliceBegin = start / kWave;
    const int nslices = (end + kWave - 1) / kWave - sliceBegin;
    const int nblocks = (nslices + kWavesPerBlock - 1) / kWavesPerBlock;
    const int grid = grid_for_blocks(nblocks);
    hipStream_t st = (hipStream_t)stream;
    if (t->layout == 2) {
        // a row range maps to slices anywhere in the locality order: every
        // slice is visited and rows outside [start, end) are not written
        TiledArgs a;
        std::memset(&a, 0, sizeof(a));
        a.srcStride = srcDesc.stride;
        a.perm = t->perm;
        a.nuniq = t->nuniq;
        a.ctrlOff = t->ctrlOff;
        a.ctrl = t->ctrl;
        a.sliceOff = t->sliceOff;
        a.sliceWidth = t->sliceWidth;
        a.rec = t->rec;
        a.rowBegin = start;
        a.rowEnd = end;
        a.nslices = t->nslices;
        a.nblocks = (t->nslices + kWavesPerBlock - 1) / kWavesPerBlock;
        const int tgrid = grid_for_blocks(a.nblocks);
        for (int e0 = 0; e0 < srcDesc.length; e0 += 8) {
            const int lc = std::min(8, srcDesc.length - e0);
            a.src = src + srcDesc.offset + e0;
            for (int k = 0; k < numOutputs; ++k) {
                a.dst[k] = outputs[k] ? outputs[k] + outputDescs[k].offset + e0 : nullptr;
                a.dstStride[k] = outputs[k] ? outputDescs[k].stride : 0;
            }
            const int g = t->group == 4 ? 4 : 8;
            if (t->nw == 1) launch_tiled<1>(lc, g, a, tgrid, st);
            else if (t->nw == 3) launch_tiled<3>(lc, g, a, tgrid, st);
            else launch_tiled<6>(lc, g, a, tgrid, st);
            int rc = last_error();
            if (rc != OSDHIP_OK) return rc;
        }
        return OSDHIP_OK;
    }
    if (t->layout == 0) {
        RecArgs a;
        std::memset(&a, 0, sizeof(a));
        a.srcStride = srcDesc.stride;
        a.rowSize = t->rowSize;
        a.sliceOff = t->sliceOff;
        a.sliceWidth = t->sliceWidth;
        a.rec = t->rec;
        a.rowBegin = start;
        a.rowEnd = end;
        a.sliceBegin = sliceBegin;
        a.nslices = nslices;
        a.nblocks = nblocks;
        for (int e0 = 0; e0 < srcDesc.length; e0 += 8) {
            const int lc = std::min(8, srcDesc.length - e0);
            a.src = src + srcDesc.offset + e0;
            for (int k = 0; k < numOutputs; ++k) {
                a.dst[k] = outputs[k] ? outputs[k] + outputDescs[k].offset + e0 : nullptr;
                a.dstStride[k] = outputs[k] ? outputDescs[k].stride : 0;
            }
            // the kernel computes every weight array of the table; only the
            // requested outputs are written
            if (t->nw == 1) launch_rec<1>(lc, t->group, a, grid, st);
            else if (t->nw == 3) launch_rec<3>(lc, t->group, a, grid, st);
            else launch_rec<6>(lc, t->group, a, grid, st);
            int rc = last_error();
            if (rc != OSDHIP_OK) return rc;
        }
        return OSDHIP_OK;
    }
    if (t->layout == 3) {
        RecArgs a;
        std::memset(&a, 0, sizeof(a));
        a.srcStride = srcDesc.stride;
        a.rowSize = t->rowSize;
        a.sliceOff = t->sliceOff;
        a.sliceWidth = t->sliceWidth;
        a.rec = t->rec;
        a.rowBegin = start;
        a.rowEnd = end;
        a.sliceBegin = sliceBegin;
        a.nslices = nslices;
        a.nblocks = nblocks;
        for (int e0 = 0; e0 < srcDesc.length; e0 += 8) {
            const int lc = std::min(8, srcDesc.length - e0);
            a.src = src + srcDesc.offset + e0;
            a.dst[0] = outputs[0] ? outputs[0] + outputDescs[0].offset + e0 : nullptr;
            a.dstStride[0] = outputs[0] ? outputDescs[0].stride : 0;
            launch_rec_paired(lc, t->group, a, grid, st);
            int rc = last_error();
            if (rc != OSDHIP_OK) return rc;
        }
        return OSDHIP_OK;
    }
    SlicedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.srcStride = srcDesc.stride;
    a.rowSize = t->rowSize;
    a.sliceOff = t->sliceOff;
    a.sliceWidth = t->sliceWidth;
    a.idx = t->sidx;
    for (int k = 0; k < numOutputs; ++k) a.w[k] = t->sw[k];
    a.rowBegin = start;
    a.rowEnd = end;
    a.sliceBegin = sliceBegin;
    a.nslices = nslices;
    a.nblocks = nblocks;
    for (int e0 = 0; e0 < srcDesc.length; e0 += 8) {
        const int lc = std::min(8, srcDesc.length - e0);
        a.src = src + srcDesc.offset + e0;
        for (int k = 0; k < numOutputs; ++k) {
            a.dst[k] = outputs[k] ? outputs[k] + outputDescs[k].offset + e0 : nullptr;
            a.dstStride[k] = outputs[k] ? outputDescs[k].stride : 0;
        }
        if (numOutputs == 1) launch_sliced<1>(lc, a, grid, st);
        else if (numOutputs == 3) launch_sliced<3>(lc, a, grid, st);
        else launch_sliced<6>(lc, a, grid, st);
        int rc = last_error();
        if (rc != OSDHIP_OK) return rc;
    }
    return OSDHIP_OK;
}

// How eval_patches orders its coords: caller order, or patch order with the
// sort's workspace from hipMallocAsync / from the caller (`ws`).
enum class PatchOrder { Caller, SortedMalloc, SortedWorkspace };

static int eval_patches(const float *src, osdhip_buffer_desc srcDesc, float *const *outs,
                        const osdhip_buffer_desc *odsIn, int numPatchCoords,
                        const osdhip_patch_coord *coords, const osdhip_patch_array *arrays,
                        const int *patchIndices, const osdhip_patch_param *params,
                        PatchOrder order, char *ws, void *stream)
{
    float *dst = outs[0], *du = outs[1], *dv = outs[2], *duu = outs[3], *duv = outs[4],
          *dvv = outs[5];
    osdhip_buffer_desc ods[6] = { odsIn[0], odsIn[1], odsIn[2], odsIn[3], odsIn[4], odsIn[5] };
    if (!src) return OSDHIP_ERR_ARGS;
    if (!outputs_ok(srcDesc, 6, outs, ods, nullptr)) return OSDHIP_ERR_ARGS;
    if (numPatchCoords <= 0) return OSDHIP_OK;
    if (!coords || !arrays || !patchIndices || !params) return OSDHIP_ERR_ARGS;

    const bool d2 = duu || duv || dvv;
    const bool d1 = d2 || du || dv;
    PatchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.srcStride = srcDesc.stride;
    a.n = numPatchCoords;
    a.coords = coords;
    a.arrays = arrays;
    a.indices = patchIndices;
    a.params = params;
    hipStream_t st = (hipStream_t)stream;

    // Patch-ordered evaluation: the coords binned by patch (patch_sort.h) in
    // stream-ordered scratch memory, then evaluation in that order.
    // Outputs still go to each coord's own row.
    void *owned = nullptr;
    if (order == PatchOrder::SortedMalloc) {
        if (hipMallocAsync(&owned, PatchSortLayout(numPatchCoords).bytes, st) == hipSuccess) {
            ws = (char *)owned;
        } else {
            (void)hipGetLastError();          // no scratch: caller order, same results
            owned = nullptr;
            ws = nullptr;
        }
    }
    if (order != PatchOrder::Caller && ws) {
        a.order = queue_patch_sort(coords, numPatchCoords, ws, st);
        int rc = last_error();
        if (rc != OSDHIP_OK) {
            if (owned) (void)hipFreeAsync(owned, st);
            return rc;
        }
    }
    int rc = OSDHIP_OK;
    for (int e0 = 0; e0 < srcDesc.length && rc == OSDHIP_OK; e0 += 8) {
        const int lc = std::min(8, srcDesc.length - e0);
        a.src = src + srcDesc.offset + e0;
        for (int o = 0; o < 6; ++o) {
            a.dst[o] = outs[o] ? outs[o] + ods[o].offset + e0 : nullptr;
            a.dstStride[o] = outs[o] ? ods[o].stride : 0;
        }
        if (d2) launch_patches<true, true>(lc, a, st);
        else if (d1) launch_patches<true, false>(lc, a, st);
        else launch_patches<false, false>(lc, a, st);
        rc = last_error();
    }
    if (owned) {
        int rf = (int)hipFreeAsync(owned, st);
        if (rc == OSDHIP_OK) rc = rf;
    }
    return rc;
}

int osdhip_eval_patches(const float *src, osdhip_buffer_desc srcDesc,
                        float *dst, osdhip_buffer_desc dstDesc,
                        float *du, osdhip_buffer_desc duDesc,
                        float *dv, osdhip_buffer_desc dvDesc,
                        float *duu, osdhip_buffer_desc duuDesc,
                        float *duv, osdhip_buffer_desc duvDesc,
                        float *dvv, osdhip_buffer_desc dvvDesc,
                        int numPatchCoords,
                        const osdhip_patch_coord *coords,
                        const osdhip_patch_array *arrays,
                        const int *patchIndices,
                        const osdhip_patch_param *params,
                        void *stream)
{
    float *outs[6] = { dst, du, dv, duu, duv, dvv };
    const osdhip_buffer_desc ods[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
    return eval_patches(src, srcDesc, outs, ods, numPatchCoords, coords, arrays, patchIndices,
                        params, PatchOrder::Caller, nullptr, stream);
}

extern "C++" {
namespace osdhip {
// Patch-ordered EvalPatches in a workspace the caller provides
// (PatchSortLayout(numPatchCoords).bytes), for osdhip_eval_patches_sorted_noalloc.
int eval_patches_in_workspace(const float *src, osdhip_buffer_desc srcDesc, float *const *outs,
                              const osdhip_buffer_desc *ods, int numPatchCoords,
                              const osdhip_patch_coord *coords, const osdhip_patch_array *arrays,
                              const int *patchIndices, const osdhip_patch_param *params,
                              char *ws, void *stream)
{
    return eval_patches(src, srcDesc, outs, ods, numPatchCoords, coords, arrays, patchIndices,
                        params, PatchOrder::SortedWorkspace, ws, stream);
}
}  // namespace osdhip
}  // extern "C++"

int osdhip_eval_patches_sorted(const float *src, osdhip_buffer_desc srcDesc,
                               float *dst, osdhip_buffer_desc dstDesc,
                               float *du, osdhip_buffer_desc duDesc,
                               float *dv, osdhip_buffer_desc dvDesc,
                               float *duu, osdhip_buffer_desc duuDesc,
                               float *duv, osdhip_buffer_desc duvDesc,
                               float *dvv, osdhip_buffer_desc dvvDesc,
                               int numPatchCoords,
                               const osdhip_patch_coord *coords,
                               const osdhip_patch_array *arrays,
                               const int *patchIndices,
                               const osdhip_patch_param *params,
                               void *stream)
{
    float *outs[6] = { dst, du, dv, duu, duv, dvv };
    const osdhip_buffer_desc ods[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
    return eval_patches(src, srcDesc, outs, ods, numPatchCoords, coords, arrays, patchIndices,
                        params, PatchOrder::SortedMalloc, nullptr, stream);
}

struct osdhip_patch_map {
    osdhip::v4i *nodes = nullptr;
    int *handles = nullptr;
    int minFace = 0, maxFace = -1, maxDepth = 0;
};

void osdhip_patch_map_destroy(osdhip_patch_map *map)
{
    if (!map) return;
    if (map->nodes) (void)hipFree(map->nodes);
    if (map->handles) (void)hipFree(map->handles);
    delete map;
}

osdhip_patch_map *osdhip_patch_map_create(const osdhip_patch_array *arrays, int numArrays,
                                          const osdhip_patch_param *params, int numPatches,
                                          void *stream)
{
    try {
        if (!arrays || !params || numArrays <= 0 || numPatches <= 0) return nullptr;
        // handles: far/patchMap.cpp:100-125 (patches in array order, vertIndex
        // counted within the array)
        std::vector<int> handles(3 * (size_t)numPatches, -1);
        long long total = 0;
        for (int a = 0; a < numArrays; ++a) {
            const osdhip_patch_array &A = arrays[a];
            if (A.numPatches < 0 || A.stride < 0 || A.primitiveIdBase != total) return nullptr;
            total += A.numPatches;
            if (total > numPatches) return nullptr;
            for (int j = 0; j < A.numPatches; ++j) {
                const size_t p = (size_t)A.primitiveIdBase + (size_t)j;
                handles[3 * p] = a;
                handles[3 * p + 1] = (int)p;
                handles[3 * p + 2] = j * A.stride;
            }
        }
        if (total != numPatches) return nullptr;
        // quadtree: far/patchMap.cpp:130-200, one root per face in
        // [minFace, maxFace]; a patch at its root depth (0, or 1 under an
        // n-gon) fills all four quadrants of the root
        int minFace = (int)(params[0].field0 & 0xfffffffu), maxFace = minFace;
        for (int p = 1; p < numPatches; ++p) {
            const int f = (int)(params[p].field0 & 0xfffffffu);
            minFace = std::min(minFace, f);
            maxFace = std::max(maxFace, f);
        }
        const osdhip::v4i empty = { osdhip::kMapEmpty, osdhip::kMapEmpty, osdhip::kMapEmpty,
                                    osdhip::kMapEmpty };
        std::vector<osdhip::v4i> nodes((size_t)(maxFace - minFace) + 1, empty);
        int maxDepth = 0;
        for (int p = 0; p < numPatches; ++p) {
            const unsigned f1 = params[p].field1;
            const int depth = (int)(f1 & 0xfu), rootDepth = (int)((f1 >> 4) & 1u);
            if (depth < rootDepth) return nullptr;
            maxDepth = std::max(maxDepth, depth);
            int node = (int)(params[p].field0 & 0xfffffffu) - minFace;
            const int leaf = -(p + 1);
            if (depth == rootDepth) {
                nodes[(size_t)node] = osdhip::v4i{ leaf, leaf, leaf, leaf };
                continue;
            }
            const int u = (int)((f1 >> 22) & 0x3ffu), v = (int)((f1 >> 12) & 0x3ffu);
            for (int j = rootDepth + 1; j <= depth; ++j) {
                const int q = (((v >> (depth - j)) & 1) << 1) | ((u >> (depth - j)) & 1);
                if (j == depth) {
                    nodes[(size_t)node][q] = leaf;
                } else {
                    int c = nodes[(size_t)node][q];
                    if (c < 0) {
                        c = (int)nodes.size();
                        nodes[(size_t)node][q] = c;
                        nodes.push_back(empty);
                    }
                    node = c;
                }
            }
        }
        osdhip_patch_map *m = new (std::nothrow) osdhip_patch_map;
        if (!m) return nullptr;
        m->minFace = minFace;
        m->maxFace = maxFace;
        m->maxDepth = maxDepth;
        hipStream_t st = (hipStream_t)stream;
        const bool ok =
            hipMalloc((void **)&m->nodes, nodes.size() * sizeof(osdhip::v4i)) == hipSuccess &&
            hipMalloc((void **)&m->handles, handles.size() * sizeof(int)) == hipSuccess &&
            hipMemcpyAsync(m->nodes, nodes.data(), nodes.size() * sizeof(osdhip::v4i),
                           hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(m->handles, handles.data(), handles.size() * sizeof(int),
                           hipMemcpyHostToDevice, st) == hipSuccess &&
            hipStreamSynchronize(st) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            osdhip_patch_map_destroy(m);
            return nullptr;
        }
        return m;
    } catch (...) {
        return nullptr;
    }
}

int osdhip_patch_map_find(const osdhip_patch_map *map, int n, const int *ptexFaces,
                          const double *u, const double *v, osdhip_patch_coord *coords,
                          void *stream)
{
    if (!map || n < 0) return OSDHIP_ERR_ARGS;
    if (n == 0) return OSDHIP_OK;
    if (!ptexFaces || !u || !v || !coords) return OSDHIP_ERR_ARGS;
    osdhip::PatchMapArgs a;
    a.nodes = map->nodes;
    a.handles = map->handles;
    a.minFace = map->minFace;
    a.maxFace = map->maxFace;
    a.maxDepth = map->maxDepth;
    a.n = n;
    a.ptex = ptexFaces;
    a.u = u;
    a.v = v;
    a.coords = coords;
    const int grid = (int)(((long long)n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(osdhip::k_patch_map_find, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream,
                       a);
    return last_error();
}

int osdhip_patch_map_find_faces(const osdhip_patch_map *map, int n, const int *faces,
                                const double *u, const double *v, const int *faceSizes,
                                const int *ptexBase, int numFaces, osdhip_patch_coord *coords,
                                void *stream)
{
    if (!map || n < 0 || numFaces < 0) return OSDHIP_ERR_ARGS;
    if (n == 0) return OSDHIP_OK;
    if (!faces || !u || !v || !coords || !faceSizes || !ptexBase) return OSDHIP_ERR_ARGS;
    osdhip::FaceQueryArgs a;
    a.map.nodes = map->nodes;
    a.map.handles = map->handles;
    a.map.minFace = map->minFace;
    a.map.maxFace = map->maxFace;
    a.map.maxDepth = map->maxDepth;
    a.map.n = n;
    a.map.ptex = nullptr;
    a.map.u = u;
    a.map.v = v;
    a.map.coords = coords;
    a.faces = faces;
    a.faceSizes = faceSizes;
    a.ptexBase = ptexBase;
    a.numFaces = numFaces;
    const int grid = (int)(((long long)n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(osdhip::k_patch_map_find_faces, dim3(grid), dim3(kBlock), 0,
                       (hipStream_t)stream, a);
    return last_error();
}

int osdhip_tess_count(int scheme, int faceSize, int numRates, const int *rates)
{
    try {
        osdhip::TessPattern t;
        return t.build(scheme, faceSize, numRates, rates) ? t.count : OSDHIP_ERR_ARGS;
    } catch (...) {
        return OSDHIP_ERR_ARGS;
    }
}

int osdhip_tess_coords(int scheme, int faceSize, int numRates, const int *rates, int numFaces,
                       float *uv, const int *faces, const osdhip_patch_map *map,
                       const int *faceSizes, const int *ptexBase, int numBaseFaces,
                       osdhip_patch_coord *coords, void *stream)
{
    try {
        if (numFaces < 0) return OSDHIP_ERR_ARGS;
        osdhip::TessPattern t;
        if (!t.build(scheme, faceSize, numRates, rates)) return OSDHIP_ERR_ARGS;
        if (numFaces == 0) return OSDHIP_OK;
        const bool toCoords = map != nullptr;
        if ((!uv && !toCoords) || (toCoords && (!coords || !faces || !faceSizes || !ptexBase ||
                                                 numBaseFaces <= 0)) ||
            (!toCoords && coords))
            return OSDHIP_ERR_ARGS;
        if ((long long)t.count * numFaces > 0x7fffffffLL) return OSDHIP_ERR_ARGS;
        hipStream_t st = (hipStream_t)stream;
        osdhip::TessRun *druns = nullptr;
        const size_t bytes = t.runs.size() * sizeof(osdhip::TessRun);
        if (hipMallocAsync((void **)&druns, bytes, st) != hipSuccess) return last_error();
        // pageable source: the copy has read it when the call returns
        hipError_t e = hipMemcpyAsync(druns, t.runs.data(), bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            osdhip::TessRunArgs a;
            a.runs = druns;
            a.numRuns = (int)t.runs.size();
            a.perFace = t.count;
            a.numFaces = numFaces;
            a.uv = uv;
            a.q.map.nodes = toCoords ? map->nodes : nullptr;
            a.q.map.handles = toCoords ? map->handles : nullptr;
            a.q.map.minFace = toCoords ? map->minFace : 0;
            a.q.map.maxFace = toCoords ? map->maxFace : -1;
            a.q.map.maxDepth = toCoords ? map->maxDepth : 0;
            a.q.map.n = 0;
            a.q.map.ptex = nullptr;
            a.q.map.u = a.q.map.v = nullptr;
            a.q.map.coords = coords;
            a.q.faces = faces;
            a.q.faceSizes = faceSizes;
            a.q.ptexBase = ptexBase;
            a.q.numFaces = numBaseFaces;
            const long long threads = (long long)numFaces * a.numRuns;
            const int grid = (int)((threads + kBlock - 1) / kBlock);
            hipLaunchKernelGGL(osdhip::k_tess_runs, dim3(grid), dim3(kBlock), 0, st, a);
            e = hipGetLastError();
        }
        const hipError_t f = hipFreeAsync(druns, st);
        if (e != hipSuccess) return (int)e;
        return f == hipSuccess ? OSDHIP_OK : (int)f;
    } catch (...) {
        return OSDHIP_ERR_ARGS;
    }
}

int osdhip_tess_uniform_quads(int numFaces, const int *ptexFaces, int rate, float *uv,
                              const osdhip_patch_map *map, osdhip_patch_coord *coords,
                              void *stream)
{
    if (numFaces < 0 || rate < 1 || rate > 4096) return OSDHIP_ERR_ARGS;
    if (numFaces == 0) return OSDHIP_OK;
    if ((map && (!coords || !ptexFaces)) || (!map && coords) || (!uv && !map))
        return OSDHIP_ERR_ARGS;
    const long long perFace = (long long)(rate + 1) * (long long)(rate + 1);
    const long long total = perFace * (long long)numFaces;
    if (total > 0x7fffffffLL) return OSDHIP_ERR_ARGS;
    osdhip::TessArgs a;
    a.numFaces = numFaces;
    a.rate = rate;
    a.perFace = (int)perFace;
    a.ptex = ptexFaces;
    a.uv = uv;
    a.map.nodes = map ? map->nodes : nullptr;
    a.map.handles = map ? map->handles : nullptr;
    a.map.minFace = map ? map->minFace : 0;
    a.map.maxFace = map ? map->maxFace : -1;
    a.map.maxDepth = map ? map->maxDepth : 0;
    a.map.n = (int)total;
    a.map.ptex = nullptr;
    a.map.u = a.map.v = nullptr;
    a.map.coords = coords;
    const int grid = (int)((total + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(osdhip::k_tess_uniform_quads, dim3(grid), dim3(kBlock), 0,
                       (hipStream_t)stream, a);
    return last_error();
}

int osdhip_limit_normals(const float *dPdu, osdhip_buffer_desc duDesc, const float *dPdv,
                         osdhip_buffer_desc dvDesc, float *normals, osdhip_buffer_desc nDesc,
                         int count, int normalize, void *stream)
{
    if (count <= 0) return OSDHIP_OK;
    if (!dPdu || !dPdv || !normals || duDesc.length < 3 || dvDesc.length < 3 || nDesc.length < 3 ||
        !desc_ok(duDesc) || !desc_ok(dvDesc) || !desc_ok(nDesc))
        return OSDHIP_ERR_ARGS;
    const int grid = (int)(((long long)count + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_limit_normals, dim3(grid), dim3(kBlock), 0, (hipStream_t)stream, dPdu,
                       duDesc.offset, duDesc.stride, dPdv, dvDesc.offset, dvDesc.stride, normals,
                       nDesc.offset, nDesc.stride, count, normalize ? 1 : 0);
    return last_error();
}

int osdhip_malloc(void **ptr, size_t bytes)
{
    if (!ptr) return OSDHIP_ERR_ARGS;
    *ptr = nullptr;
    if (bytes == 0) return OSDHIP_OK;
    return (int)hipMalloc(ptr, bytes);
}

int osdhip_free(void *ptr) { return ptr ? (int)hipFree(ptr) : OSDHIP_OK; }

int osdhip_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return OSDHIP_OK;
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
}

int osdhip_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return OSDHIP_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e != hipSuccess) return (int)e;
    return (int)hipStreamSynchronize((hipStream_t)stream);
}

int osdhip_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return OSDHIP_OK;
    return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
}

int osdhip_stream_synchronize(void *stream) { return (int)hipStreamSynchronize((hipStream_t)stream); }

int osdhip_device_synchronize(void) { return (int)hipDeviceSynchronize(); }

int osdhip_device_info(char *buf, int buflen)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        if (buf && buflen > 0) buf[0] = 0;
        return -1;
    }
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return -1;
    if (buf && buflen > 0) {
        std::strncpy(buf, p.name, (size_t)buflen - 1);
        buf[buflen - 1] = 0;
        size_t l = std::strlen(buf);
        if (l + 2 < (size_t)buflen) {
            buf[l] = ' ';
            std::strncpy(buf + l + 1, p.gcnArchName, (size_t)buflen - l - 2);
            buf[buflen - 1] = 0;
        }
    }
    return n;
}

const char *osdhip_version(void) { return "osdhip 0.2 gfx950"; }

}  // extern "C"
