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

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/osdhip.h"
#include "common.h"
#include "patch_kernels.h"
#include "stencil_csr.h"
#include "stencil_sliced.h"

namespace osdhip {

// Tuning knobs (osdhip_config).
static int g_stencil_kernel = 0;     // table eval: 0 sliced, 1 CSR

template <int NW, int CAP, bool VEC>
static void launch_csr(int lc, const StencilArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_stencils_csr<NW, 1, CAP, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_stencils_csr<NW, 2, CAP, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_stencils_csr<NW, 3, CAP, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_stencils_csr<NW, 4, CAP, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_stencils_csr<NW, 5, CAP, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_stencils_csr<NW, 6, CAP, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_stencils_csr<NW, 7, CAP, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: hipLaunchKernelGGL((k_stencils_csr<NW, 8, CAP, VEC>), dim3(grid), dim3(kBlock), 0, st, a); break;
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

static int last_error()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? OSDHIP_OK : (int)e;
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
        // SLOTS = CAP + 4 entries per wave: 32 KB per 256-thread block for
        // NW = 1, i.e. five blocks per 160 KB CU
        constexpr int CAP = NW == 6 ? 508 : 1020;
        if (vec) launch_csr<NW, CAP, true>(lc, a, grid, st);
        else launch_csr<NW, CAP, false>(lc, a, grid, st);
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

// Device copy of a (limit) stencil table: the Far CSR arrays (for the raw
// getters) plus the 64-row sliced layout the table kernel streams.
struct osdhip_stencil_table {
    int n = 0, nw = 0, nslices = 0;
    long long nnz = 0, sliceEntries = 0, bytes = 0;
    int *sizes = nullptr, *offsets = nullptr, *indices = nullptr;
    float *w[6] = {};
    long long *sliceOff = nullptr;
    int *sliceWidth = nullptr;
    int *sidx = nullptr;
    float *sw[6] = {};
};

static void free_table(osdhip_stencil_table *t)
{
    if (!t) return;
    void *ptrs[] = { t->sizes, t->offsets, t->indices, t->w[0], t->w[1], t->w[2], t->w[3],
                     t->w[4], t->w[5], t->sliceOff, t->sliceWidth, t->sidx, t->sw[0], t->sw[1],
                     t->sw[2], t->sw[3], t->sw[4], t->sw[5] };
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete t;
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

osdhip_stencil_table *osdhip_stencil_table_create(int numStencils, const int *sizes,
                                                  const int *offsets, const int *indices,
                                                  int numWeightArrays,
                                                  const float *const *weights, void *stream)
{
    if (numStencils < 0 || (numStencils > 0 && (!sizes || !indices || !weights)) ||
        (numWeightArrays != 1 && numWeightArrays != 3 && numWeightArrays != 6))
        return nullptr;
    for (int k = 0; k < numWeightArrays; ++k)
        if (numStencils > 0 && !weights[k]) return nullptr;
    osdhip_stencil_table *t = new (std::nothrow) osdhip_stencil_table;
    if (!t) return nullptr;
    t->n = numStencils;
    t->nw = numWeightArrays;
    hipStream_t st = (hipStream_t)stream;

    // CSR mirror (offsets regenerated when absent: exclusive prefix of sizes)
    std::vector<int> off(numStencils);
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
              upload(&t->offsets, off.data(), (size_t)numStencils, st, t->bytes) &&
              upload(&t->indices, indices, (size_t)nnz, st, t->bytes);
    for (int k = 0; ok && k < numWeightArrays; ++k)
        ok = upload(&t->w[k], weights[k], (size_t)nnz, st, t->bytes);

    // sliced layout: slice s holds rows [64s, 64s+64), column-major, padded
    const int S = (numStencils + kWave - 1) / kWave;
    t->nslices = S;
    std::vector<long long> sOff(S + 1, 0);
    std::vector<int> sW(S, 0);
    for (int s = 0; s < S; ++s) {
        int w = 0;
        for (int r = s * kWave; r < std::min(numStencils, (s + 1) * kWave); ++r) w = std::max(w, sizes[r]);
        sW[s] = w;
        sOff[s + 1] = sOff[s] + (long long)w * kWave;
    }
    t->sliceEntries = sOff[S];
    std::vector<int> pidx(ok ? (size_t)t->sliceEntries : 0, 0);
    std::vector<float> pw;
    if (ok) {
        for (int s = 0; s < S; ++s)
            for (int r = s * kWave; r < std::min(numStencils, (s + 1) * kWave); ++r)
                for (int j = 0; j < sizes[r]; ++j)
                    pidx[sOff[s] + (long long)j * kWave + (r - s * kWave)] = indices[off[r] + j];
        ok = upload(&t->sliceOff, sOff.data(), (size_t)S, st, t->bytes) &&
             upload(&t->sliceWidth, sW.data(), (size_t)S, st, t->bytes) &&
             upload(&t->sidx, pidx.data(), pidx.size(), st, t->bytes) &&
             hipStreamSynchronize(st) == hipSuccess;
    }
    for (int k = 0; ok && k < numWeightArrays; ++k) {
        pw.assign((size_t)t->sliceEntries, 0.0f);
        for (int s = 0; s < S; ++s)
            for (int r = s * kWave; r < std::min(numStencils, (s + 1) * kWave); ++r)
                for (int j = 0; j < sizes[r]; ++j)
                    pw[sOff[s] + (long long)j * kWave + (r - s * kWave)] = weights[k][off[r] + j];
        ok = upload(&t->sw[k], pw.data(), pw.size(), st, t->bytes) &&
             hipStreamSynchronize(st) == hipSuccess;
    }
    if (!ok) { free_table(t); return nullptr; }
    return t;
}

void osdhip_stencil_table_destroy(osdhip_stencil_table *t) { free_table(t); }

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
    if (g_stencil_kernel == 1) {
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
    SlicedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.srcStride = srcDesc.stride;
    a.rowSize = t->sizes;
    a.sliceOff = t->sliceOff;
    a.sliceWidth = t->sliceWidth;
    a.idx = t->sidx;
    for (int k = 0; k < numOutputs; ++k) a.w[k] = t->sw[k];
    a.rowBegin = start;
    a.rowEnd = end;
    a.sliceBegin = start / kWave;
    a.nslices = (end + kWave - 1) / kWave - a.sliceBegin;
    a.nblocks = (a.nslices + kWavesPerBlock - 1) / kWavesPerBlock;
    const int grid = grid_for_blocks(a.nblocks);
    hipStream_t st = (hipStream_t)stream;
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
    if (!src) return OSDHIP_ERR_ARGS;
    float *outs[6] = { dst, du, dv, duu, duv, dvv };
    osdhip_buffer_desc ods[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
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
    for (int e0 = 0; e0 < srcDesc.length; e0 += 8) {
        const int lc = std::min(8, srcDesc.length - e0);
        a.src = src + srcDesc.offset + e0;
        for (int o = 0; o < 6; ++o) {
            a.dst[o] = outs[o] ? outs[o] + ods[o].offset + e0 : nullptr;
            a.dstStride[o] = outs[o] ? ods[o].stride : 0;
        }
        if (d2) launch_patches<true, true>(lc, a, st);
        else if (d1) launch_patches<true, false>(lc, a, st);
        else launch_patches<false, false>(lc, a, st);
        int rc = last_error();
        if (rc != OSDHIP_OK) return rc;
    }
    return OSDHIP_OK;
}

int osdhip_config(const char *key, int value)
{
    if (!key) return OSDHIP_ERR_ARGS;
    if (std::strcmp(key, "stencil_kernel") == 0) {
        if (value != 0 && value != 1) return OSDHIP_ERR_ARGS;
        int old = g_stencil_kernel;
        g_stencil_kernel = value;
        return old;
    }
    return OSDHIP_ERR_UNSUPPORTED;
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
