// EvalPatches: limit evaluation at PatchCoords (osd/cpuEvaluator.cpp:
// 157-381).  One lane per PatchCoord: PatchParam decode and Normalize, the
// B-spline / Gregory / bilinear basis in registers (patch_basis_device.h),
// then the 16 / 20 / 4-point gather for P and each requested derivative.
// Sums accumulate in the reference's point order without FMA; derivative
// sums are scaled by 2^depth / 4^depth (exact powers of two).
#pragma once
#include "../../include/osdhip.h"
#include "common.h"
#include "patch_basis_device.h"

namespace osdhip {

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
    const int *order;       // evaluation order (coord index per lane), NULL = identity
};

// acc[o][k] += v[k] * weight_o[j], in the reference's point order.
template <bool D1, bool D2, int LC, int NO>
__device__ __forceinline__ void accumulate_point(const float (&v)[LC], const Basis<D1, D2> &B,
                                                 int j, float (&acc)[NO][LC])
{
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

// The first NA = min(NP, 16) points' control values are loaded first (all
// NA loads in flight), then the basis (`eval` fills B), then the products
// are summed in point order, then the remaining points' (as the Gregory
// path does with its last 4, which bounds the registers held at once).
template <int NP, int LC, bool D1, bool D2, int NO, class EVAL>
__device__ __forceinline__ void gather_points(const float *src, size_t sstride, const int *cvs,
                                              Basis<D1, D2> &B, float (&acc)[NO][LC], EVAL eval)
{
    constexpr int NA = NP < 16 ? NP : 16;
    float v[NA][LC];
#pragma unroll
    for (int j = 0; j < NA; ++j) {
        const float *p = src + (size_t)max(cvs[j], 0) * sstride;
#pragma unroll
        for (int k = 0; k < LC; ++k) v[j][k] = p[k];
    }
    eval();
#pragma unroll
    for (int j = 0; j < NA; ++j) accumulate_point<D1, D2, LC>(v[j], B, j, acc);
    if constexpr (NP > NA) {
        float r[NP - NA][LC];
#pragma unroll
        for (int j = NA; j < NP; ++j) {
            const float *p = src + (size_t)max(cvs[j], 0) * sstride;
#pragma unroll
            for (int k = 0; k < LC; ++k) r[j - NA][k] = p[k];
        }
#pragma unroll
        for (int j = NA; j < NP; ++j) accumulate_point<D1, D2, LC>(r[j - NA], B, j, acc);
    }
}

template <bool D1, bool D2, int LC>
__global__ __launch_bounds__(kBlock) void k_patches(PatchArgs a)
{
    const int lane_i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (lane_i >= a.n) return;
    // Each coord's result depends on that coord only, so any order gives the
    // same bits; a patch-sorted order makes neighbouring lanes share control
    // points and index rows in L2.
    const int i = a.order ? a.order[lane_i] : lane_i;

    const osdhip_patch_coord c = a.coords[i];
    const osdhip_patch_array arr = a.arrays[c.arrayIndex];
    const osdhip_patch_param pp = a.params[c.patchIndex];
    const unsigned f1 = pp.field1;
    const int type = ((f1 >> 5) & 1u) ? arr.regDesc : arr.desc;
    const int boundary = (int)((f1 >> 7) & 0x1fu);

    // PatchParam Normalize / NormalizeTriangle: ptex-face (s,t) ->
    // patch-local (s,t), and the signed 2^depth derivative factor
    float s = c.s, t = c.t, sign;
    const float k1 = patch_normalize(type, f1, s, t, sign);

    // A negative point index (INDEX_INVALID of a patch built without
    // fillMissingBoundaryPoints, far/patchBuilder.h:177-183) reads point 0
    // instead of memory before the buffer; such points carry zero weight
    // after the boundary fold.
    // The point count follows from the type alone (REGULAR 16, GREGORY_BASIS
    // 20, QUADS 4, LOOP 12, GREGORY_TRIANGLE 18, TRIANGLES 3, anything else
    // 0 -> zero outputs, as the reference's Clear() without AddWithWeight),
    // so the index and control-point loads are issued before the basis is
    // evaluated and overlap with it.
    const int n = patch_points(type);
    const int *cvs = a.indices + arr.indexBase + arr.stride * (c.patchIndex - arr.primitiveIdBase);
    constexpr int NO = 1 + (D1 ? 2 : 0) + (D2 ? 3 : 0);
    float acc[NO][LC];
#pragma unroll
    for (int o = 0; o < NO; ++o)
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[o][k] = 0.0f;

    const size_t sstride = (size_t)a.srcStride;
    Basis<D1, D2> B;
    if (n == 16 || n == 20) {
        float v[16][LC];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float *p = a.src + (size_t)max(cvs[j], 0) * sstride;
#pragma unroll
            for (int k = 0; k < LC; ++k) v[j][k] = p[k];
        }
        if (n == 16) basis_bspline<D1, D2>(s, t, boundary, B);
        else basis_gregory<D1, D2>(s, t, B);
#pragma unroll
        for (int j = 0; j < 16; ++j) accumulate_point<D1, D2, LC>(v[j], B, j, acc);
        if (n == 20) {
            float v4[4][LC];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float *p = a.src + (size_t)max(cvs[16 + j], 0) * sstride;
#pragma unroll
                for (int k = 0; k < LC; ++k) v4[j][k] = p[k];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) accumulate_point<D1, D2, LC>(v4[j], B, 16 + j, acc);
        }
    } else if (n == 4) {
        float v[4][LC];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float *p = a.src + (size_t)max(cvs[j], 0) * sstride;
#pragma unroll
            for (int k = 0; k < LC; ++k) v[j][k] = p[k];
        }
        basis_linear<D1, D2>(s, t, B);
#pragma unroll
        for (int j = 0; j < 4; ++j) accumulate_point<D1, D2, LC>(v[j], B, j, acc);
    } else if (n == 12) {
        gather_points<12, LC>(a.src, sstride, cvs, B, acc,
                              [&] { basis_box_tri<D1, D2>(s, t, boundary, B); });
    } else if (n == 18) {
        gather_points<18, LC>(a.src, sstride, cvs, B, acc,
                              [&] { basis_gregory_tri<D1, D2>(s, t, B); });
    } else if (n == 3) {
        gather_points<3, LC>(a.src, sstride, cvs, B, acc,
                             [&] { basis_linear_tri<D1, D2>(s, t, B); });
    }
    // +-2^depth (first) and +-4^depth (second derivatives): exact signed
    // power-of-two factors, applied to the sums instead of to every weight
    // (the second-derivative factor is sign * k1 * k1, as the reference's).
    const float k2 = sign * k1 * k1;
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

// Limit normals from EvalPatches' first derivatives, one thread per point:
// n = cross(du, dv) component by component as written (no fused
// multiply-add), then n * (1 / sqrtf(dot)) when normalizing (IEEE division
// and square root), which are the reference tutorial's steps.
__global__ __launch_bounds__(kBlock) void k_limit_normals(const float *du, int duOff, int duStride,
                                                          const float *dv, int dvOff, int dvStride,
                                                          float *n, int nOff, int nStride,
                                                          int count, int normalize)
{
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= count) return;
    const float *a = du + duOff + (size_t)i * (size_t)duStride;
    const float *b = dv + dvOff + (size_t)i * (size_t)dvStride;
    const float a0 = a[0], a1 = a[1], a2 = a[2], b0 = b[0], b1 = b[1], b2 = b[2];
    float c0 = a1 * b2 - a2 * b1;
    float c1 = a2 * b0 - a0 * b2;
    float c2 = a0 * b1 - a1 * b0;
    if (normalize) {
        const float rn = 1.0f / sqrtf(c0 * c0 + c1 * c1 + c2 * c2);
        c0 *= rn;
        c1 *= rn;
        c2 *= rn;
    }
    float *o = n + nOff + (size_t)i * (size_t)nStride;
    o[0] = c0;
    o[1] = c1;
    o[2] = c2;
}

}  // namespace osdhip
