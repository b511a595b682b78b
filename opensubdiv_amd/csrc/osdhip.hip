// libosdhip: MI355X (gfx950, CDNA4) kernels for the Osd evaluator hot path
// and the C ABI declared in include/osdhip.h.
//
// Stencil evaluation is a sparse gather (CSR rows x a dense multi-column
// source), HBM-bound at < 0.3 FLOP/byte, so it uses no MFMA.  Each 64-lane
// wavefront owns 64 consecutive stencils: it stages the wave's contiguous
// slice of the index / weight streams into LDS with coalesced non-temporal
// loads (the streams are read exactly once), then every lane reduces its own
// stencil from LDS while gathering control points through L2.  All weight
// arrays of a derivative table (P, du, dv [, duu, duv, dvv]) share the index
// stream and are applied in ONE launch.  Work-groups are remapped so each of
// the 8 XCDs (each with its own L2) walks a contiguous range of stencils.
//
// Patch evaluation computes the B-spline / Gregory / bilinear basis in
// registers per PatchCoord and gathers the 16 / 20 / 4 control points.
//
// Parity: built with -ffp-contract=off (no fused multiply-add) and every
// sum accumulates w * x in the reference's index order starting from 0, so
// stencil and patch results equal Osd::CpuEvaluator's bit for bit.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstring>

#include "../../include/osdhip.h"
#include "patch_basis_device.h"

namespace osdhip {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kNumXcd = 8;

// Wave-scope LDS hand-off: LDS traffic of one wave is processed in order,
// so only the compiler has to be kept from reordering across this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Hardware dispatches work-group b to XCD (b % 8).  Map it to a logical
// block so that XCD x handles the contiguous range [x*per, (x+1)*per).
// gridDim.x is always a multiple of 8.
__device__ __forceinline__ int xcd_logical_block()
{
    const int per = (int)gridDim.x / kNumXcd;
    return ((int)blockIdx.x % kNumXcd) * per + (int)blockIdx.x / kNumXcd;
}

struct StencilArgs {
    const float *src;       // includes srcDesc.offset and the element window start
    int srcStride;
    float *dst[6];          // include their desc offsets (+ element window start)
    int dstStride[6];
    const int *sizes, *offsets, *indices;
    const float *w[6];
    int start, count;       // stencils [start, start+count); output row = i - start
    int nblocks;            // logical blocks that hold stencils
};

// NW weight arrays (1: P, 3: P/du/dv, 6: + duu/duv/dvv); LC floats per vertex
// handled by this launch; CAP = table entries staged per wave.
template <int NW, int LC, int CAP>
__global__ __launch_bounds__(kBlock) void k_stencils(StencilArgs a)
{
    __shared__ int s_idx[kWavesPerBlock][CAP];
    __shared__ float s_w[kWavesPerBlock][NW][CAP];

    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wv = (int)threadIdx.x / kWave;
    const int lb = xcd_logical_block();
    if (lb >= a.nblocks) return;
    const int first = lb * kBlock + wv * kWave;     // relative to a.start
    if (first >= a.count) return;                    // whole wave idle
    const int cnt = min(kWave, a.count - first);
    const bool active = lane < cnt;
    const int i = a.start + first + lane;

    int off = 0, sz = 0;
    if (active) {
        off = a.offsets[i];
        sz = a.sizes[i];
    }
    // Table slice [lo, hi) covering every stencil of this wave.
    int lo = active ? off : INT_MAX;
    int hi = active ? off + sz : INT_MIN;
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
        lo = min(lo, __shfl_xor(lo, m));
        hi = max(hi, __shfl_xor(hi, m));
    }
    const int n = hi - lo;

    float acc[NW][LC];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[w][k] = 0.0f;

    const float *__restrict__ src = a.src;
    const size_t sstride = (size_t)a.srcStride;

    if (n <= CAP) {
        // Coalesced, streaming (non-temporal) copy of the slice into LDS.
        for (int e = lane; e < n; e += kWave) {
            s_idx[wv][e] = __builtin_nontemporal_load(a.indices + lo + e);
#pragma unroll
            for (int w = 0; w < NW; ++w)
                s_w[wv][w][e] = __builtin_nontemporal_load(a.w[w] + lo + e);
        }
        wave_lds_sync();
        const int base = off - lo;
        for (int j = 0; j < sz; ++j) {
            const int idx = s_idx[wv][base + j];
            const float *p = src + (size_t)(unsigned)idx * sstride;
            float v[LC];
#pragma unroll
            for (int k = 0; k < LC; ++k) v[k] = p[k];
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const float ww = s_w[wv][w][base + j];
#pragma unroll
                for (int k = 0; k < LC; ++k) acc[w][k] += v[k] * ww;
            }
        }
    } else {
        // Slice larger than the staging buffer: read the table directly.
        for (int j = 0; j < sz; ++j) {
            const int idx = a.indices[off + j];
            const float *p = src + (size_t)(unsigned)idx * sstride;
            float v[LC];
#pragma unroll
            for (int k = 0; k < LC; ++k) v[k] = p[k];
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const float ww = a.w[w][off + j];
#pragma unroll
                for (int k = 0; k < LC; ++k) acc[w][k] += v[k] * ww;
            }
        }
    }

    if (active) {
        const size_t row = (size_t)(first + lane);
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            float *d = a.dst[w];
            if (!d) continue;
            d += row * (size_t)a.dstStride[w];
#pragma unroll
            for (int k = 0; k < LC; ++k) d[k] = acc[w][k];
        }
    }
}

struct PatchArgs {
    const float *src;       // includes srcDesc.offset and the element window start
    int srcStride;
    float *dst[6];          // P, du, dv, duu, duv, dvv (NULL = skip)
    int dstStride[6];
    int n;
    const osdhip_patch_coord *coords;
    const osdhip_patch_array *arrays;
    const int *indices;
    const osdhip_patch_param *params;
};

template <bool D1, bool D2, int LC>
__global__ __launch_bounds__(kBlock) void k_patches(PatchArgs a)
{
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= a.n) return;

    const osdhip_patch_coord c = a.coords[i];
    const osdhip_patch_array arr = a.arrays[c.arrayIndex];
    const osdhip_patch_param pp = a.params[c.patchIndex];
    const unsigned f1 = pp.field1;
    const int type = ((f1 >> 5) & 1u) ? arr.regDesc : arr.desc;
    const int depth = (int)(f1 & 0xfu);
    const int nonquad = (int)((f1 >> 4) & 1u);
    const int boundary = (int)((f1 >> 7) & 0x1fu);
    const int pv = (int)((f1 >> 12) & 0x3ffu);
    const int pu = (int)((f1 >> 22) & 0x3ffu);

    // PatchParam::Normalize: ptex-face (s,t) -> patch-local (s,t)
    const float frac_inv = (float)(1 << (depth - nonquad));
    const float s = c.s * frac_inv - (float)pu;
    const float t = c.t * frac_inv - (float)pv;

    Basis<D1, D2> B;
    B.n = 0;
    if (type == OSDHIP_PATCH_REGULAR) basis_bspline<D1, D2>(s, t, boundary, B);
    else if (type == OSDHIP_PATCH_GREGORY_BASIS) basis_gregory<D1, D2>(s, t, B);
    else if (type == OSDHIP_PATCH_QUADS) basis_linear<D1, D2>(s, t, B);

    const int *cvs = a.indices + arr.indexBase + arr.stride * (c.patchIndex - arr.primitiveIdBase);
    constexpr int NO = 1 + (D1 ? 2 : 0) + (D2 ? 3 : 0);
    float acc[NO][LC];
#pragma unroll
    for (int o = 0; o < NO; ++o)
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[o][k] = 0.0f;

    const size_t sstride = (size_t)a.srcStride;
#pragma unroll
    for (int j = 0; j < 20; ++j) {
        if (j < B.n) {
            const float *p = a.src + (size_t)(unsigned)cvs[j] * sstride;
            float v[LC];
#pragma unroll
            for (int k = 0; k < LC; ++k) v[k] = p[k];
#pragma unroll
            for (int k = 0; k < LC; ++k) {
                acc[0][k] += v[k] * B.p[j];
                if constexpr (D1) {
                    acc[1][k] += v[k] * B.ds[j];
                    acc[2][k] += v[k] * B.dt[j];
                }
                if constexpr (D2) {
                    acc[3][k] += v[k] * B.dss[j];
                    acc[4][k] += v[k] * B.dst[j];
                    acc[5][k] += v[k] * B.dtt[j];
                }
            }
        }
    }
    // 2^depth (first) and 4^depth (second derivatives): exact power-of-two
    // factors, applied to the sums instead of to every weight.
    const float k1 = (float)(1 << depth);
    const float k2 = k1 * k1;
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        float *d = a.dst[o];
        if (!d) continue;
        const float f = o == 0 ? 1.0f : (o < 3 ? k1 : k2);
        d += (size_t)i * (size_t)a.dstStride[o];
#pragma unroll
        for (int k = 0; k < LC; ++k) d[k] = acc[o][k] * f;
    }
}

// ---------------------------------------------------------------------------
// host-side dispatch

template <int NW, int CAP>
static int launch_stencils_lc(int lc, const StencilArgs &a, int grid, hipStream_t st)
{
    switch (lc) {
    case 1: hipLaunchKernelGGL((k_stencils<NW, 1, CAP>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_stencils<NW, 2, CAP>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_stencils<NW, 3, CAP>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_stencils<NW, 4, CAP>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_stencils<NW, 5, CAP>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_stencils<NW, 6, CAP>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 7: hipLaunchKernelGGL((k_stencils<NW, 7, CAP>), dim3(grid), dim3(kBlock), 0, st, a); break;
    case 8: hipLaunchKernelGGL((k_stencils<NW, 8, CAP>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: return OSDHIP_ERR_UNSUPPORTED;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? OSDHIP_OK : (int)e;
}

static bool desc_ok(const osdhip_buffer_desc &d)
{
    return d.length > 0 && d.stride > 0 && d.offset >= 0 &&
           d.length <= d.stride - (d.offset % d.stride);
}

// NW outputs, outs[w]/ods[w] may have NULL pointers (skipped).
template <int NW>
static int eval_stencils_impl(const float *src, osdhip_buffer_desc sd,
                              float *const *outs, const osdhip_buffer_desc *ods,
                              const int *sizes, const int *offsets,
                              const int *indices, const float *const *ws,
                              int start, int end, void *stream)
{
    if (end <= start) return OSDHIP_OK;
    if (!src || !sizes || !offsets || !indices || start < 0) return OSDHIP_ERR_ARGS;
    if (!desc_ok(sd)) return OSDHIP_ERR_ARGS;
    bool any = false;
    for (int w = 0; w < NW; ++w) {
        if (!outs[w]) continue;
        if (ods[w].length != sd.length || !desc_ok(ods[w]) || !ws[w]) return OSDHIP_ERR_ARGS;
        any = true;
    }
    if (!any) return OSDHIP_ERR_ARGS;

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
    const int grid = (a.nblocks + kNumXcd - 1) / kNumXcd * kNumXcd;
    hipStream_t st = (hipStream_t)stream;
    // Elements beyond 8 per vertex are handled by further launches over the
    // same stencils (windows of up to 8 floats).
    for (int e0 = 0; e0 < sd.length; e0 += 8) {
        const int lc = sd.length - e0 < 8 ? sd.length - e0 : 8;
        a.src = src + sd.offset + e0;
        for (int w = 0; w < NW; ++w) {
            a.dst[w] = outs[w] ? outs[w] + ods[w].offset + e0 : nullptr;
            a.dstStride[w] = outs[w] ? ods[w].stride : 0;
            // a skipped output still needs a readable weight stream for the
            // staging loop; reuse the P weights (their products are dropped)
            a.w[w] = ws[w] ? ws[w] : ws[0];
        }
        int rc = launch_stencils_lc<NW, NW == 6 ? 512 : 1024>(lc, a, grid, st);
        if (rc != OSDHIP_OK) return rc;
    }
    return OSDHIP_OK;
}

template <bool D1, bool D2>
static int launch_patches_lc(int lc, const PatchArgs &a, hipStream_t st)
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
    case 8: hipLaunchKernelGGL((k_patches<D1, D2, 8>), dim3(grid), dim3(kBlock), 0, st, a); break;
    default: return OSDHIP_ERR_UNSUPPORTED;
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? OSDHIP_OK : (int)e;
}

}  // namespace osdhip

using namespace osdhip;

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
    return eval_stencils_impl<1>(src, srcDesc, outs, ods, sizes, offsets, indices, ws,
                                 start, end, stream);
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
    if (!ws[0]) return OSDHIP_ERR_ARGS;
    return eval_stencils_impl<3>(src, srcDesc, outs, ods, sizes, offsets, indices, ws,
                                 start, end, stream);
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
    if (!ws[0]) return OSDHIP_ERR_ARGS;
    return eval_stencils_impl<6>(src, srcDesc, outs, ods, sizes, offsets, indices, ws,
                                 start, end, stream);
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
    if (!src || !desc_ok(srcDesc)) return OSDHIP_ERR_ARGS;
    float *outs[6] = { dst, du, dv, duu, duv, dvv };
    osdhip_buffer_desc ods[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
    bool any = false;
    for (int o = 0; o < 6; ++o) {
        if (!outs[o]) continue;
        if (ods[o].length != srcDesc.length || !desc_ok(ods[o])) return OSDHIP_ERR_ARGS;
        any = true;
    }
    if (!any) return OSDHIP_ERR_ARGS;
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
        const int lc = srcDesc.length - e0 < 8 ? srcDesc.length - e0 : 8;
        a.src = src + srcDesc.offset + e0;
        for (int o = 0; o < 6; ++o) {
            a.dst[o] = outs[o] ? outs[o] + ods[o].offset + e0 : nullptr;
            a.dstStride[o] = outs[o] ? ods[o].stride : 0;
        }
        int rc;
        if (d2) rc = launch_patches_lc<true, true>(lc, a, st);
        else if (d1) rc = launch_patches_lc<true, false>(lc, a, st);
        else rc = launch_patches_lc<false, false>(lc, a, st);
        if (rc != OSDHIP_OK) return rc;
    }
    return OSDHIP_OK;
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

const char *osdhip_version(void) { return "osdhip 0.1 gfx950"; }

}  // extern "C"
