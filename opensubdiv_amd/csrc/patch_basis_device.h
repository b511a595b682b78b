// Device-side patch basis for the limit-evaluation kernel (gfx950).
//
// Evaluates, in registers, the weights that OpenSubdiv's
// OsdEvaluatePatchBasis (osd/patchBasis.h:1447-1610) produces for the quad
// patch types: REGULAR bicubic B-spline with the boundary fold
// (patchBasis.h:141-245), GREGORY_BASIS 20-point Gregory with the default
// approximate derivatives (patchBasis.h:334-490) and QUADS bilinear
// (patchBasis.h:55-97).  Weights are produced UNSCALED; the caller applies
// the 2^depth / 4^depth derivative factors to the accumulated sums, which is
// exact because the factors are powers of two.
//
// Parity: the translation unit is compiled with -ffp-contract=off and every
// product / sum is written in the reference's order, so results match
// Osd::CpuEvaluator bit for bit.
//
// All arrays are fixed-size and indexed by compile-time constants after
// unrolling, so they live in VGPRs.
#pragma once
#include <hip/hip_runtime.h>

namespace osdhip {

struct Curve {
    float b[4], d[4], dd[4];
};

// Uniform cubic B-spline, value / first / second derivative.  The
// polynomials use the same expanded forms and operation order as
// osd/patchBasis.h:101-133 and the library is compiled without FP
// contraction, so the weights round exactly as CpuEvaluator's do.
__device__ __forceinline__ void bspline_curve(float t, Curve &c)
{
    const float k6 = 1.0f / 6.0f;
    float t2 = t * t, t3 = t * t2;
    c.b[0] = k6 * (1.0f - 3.0f * (t - t2) - t3);
    c.b[1] = k6 * (4.0f - 6.0f * t2 + 3.0f * t3);
    c.b[2] = k6 * (1.0f + 3.0f * (t + t2 - t3));
    c.b[3] = k6 * t3;
    c.d[0] = -0.5f * t2 + t - 0.5f;
    c.d[1] = 1.5f * t2 - 2.0f * t;
    c.d[2] = -1.5f * t2 + t + 0.5f;
    c.d[3] = 0.5f * t2;
    c.dd[0] = -t + 1.0f;
    c.dd[1] = 3.0f * t - 2.0f;
    c.dd[2] = -3.0f * t + 1.0f;
    c.dd[3] = t;
}

// Cubic Bernstein basis, value / first / second derivative, in the forms of
// osd/patchBasis.h:250-287.
__device__ __forceinline__ void bezier_curve(float t, Curve &c)
{
    float t2 = t * t, u = 1.0f - t, u2 = u * u;
    c.b[0] = u2 * u;
    c.b[1] = u2 * t * 3.0f;
    c.b[2] = t2 * u * 3.0f;
    c.b[3] = t2 * t;
    c.d[0] = -3.0f * u2;
    c.d[1] = 9.0f * t2 - 12.0f * t + 3.0f;
    c.d[2] = -9.0f * t2 + 6.0f * t;
    c.d[3] = 3.0f * t2;
    c.dd[0] = 6.0f * u;
    c.dd[1] = 18.0f * t - 12.0f;
    c.dd[2] = -18.0f * t + 6.0f;
    c.dd[3] = 6.0f * t;
}

// Boundary fold of a 4x4 weight set (row-major, row = t, column = s): a
// missing boundary row/column point P0 is the reflection 2*P1 - P2.
__device__ __forceinline__ void fold_boundary(int mask, float (&w)[20])
{
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if (mask & 1) { w[8 + c] -= w[c]; w[4 + c] += 2.0f * w[c]; w[c] = 0.0f; }
#pragma unroll
    for (int r = 0; r < 16; r += 4)
        if (mask & 2) { w[r + 1] -= w[r + 3]; w[r + 2] += 2.0f * w[r + 3]; w[r + 3] = 0.0f; }
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if (mask & 4) { w[4 + c] -= w[12 + c]; w[8 + c] += 2.0f * w[12 + c]; w[12 + c] = 0.0f; }
#pragma unroll
    for (int r = 0; r < 16; r += 4)
        if (mask & 8) { w[r + 2] -= w[r]; w[r + 1] += 2.0f * w[r]; w[r] = 0.0f; }
}

// Weights of one patch evaluation.  n = number of control points used.
template <bool D1, bool D2>
struct Basis {
    float p[20];
    float ds[D1 ? 20 : 1], dt[D1 ? 20 : 1];
    float dss[D2 ? 20 : 1], dst[D2 ? 20 : 1], dtt[D2 ? 20 : 1];
    int n;
};

template <bool D1, bool D2>
__device__ __forceinline__ void basis_bspline(float s, float t, int boundary, Basis<D1, D2> &B)
{
    Curve cs, ct;
    bspline_curve(s, cs);
    bspline_curve(t, ct);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            B.p[4 * i + j] = cs.b[j] * ct.b[i];
            if constexpr (D1) {
                B.ds[4 * i + j] = cs.d[j] * ct.b[i];
                B.dt[4 * i + j] = cs.b[j] * ct.d[i];
            }
            if constexpr (D2) {
                B.dss[4 * i + j] = cs.dd[j] * ct.b[i];
                B.dst[4 * i + j] = cs.d[j] * ct.d[i];
                B.dtt[4 * i + j] = cs.b[j] * ct.dd[i];
            }
        }
    if (boundary) {
        fold_boundary(boundary, B.p);
        if constexpr (D1) { fold_boundary(boundary, B.ds); fold_boundary(boundary, B.dt); }
        if constexpr (D2) {
            fold_boundary(boundary, B.dss);
            fold_boundary(boundary, B.dst);
            fold_boundary(boundary, B.dtt);
        }
    }
    B.n = 16;
}

// Gregory point k of corner c is 5c+k: 0 corner, 1/2 edge points, 3/4 the
// two interior face points.  Boundary points carry plain Bezier weights of
// (column, row); each interior pair blends into one Bezier interior point
// with rational factors g and 1-g.
template <bool D1, bool D2>
__device__ __forceinline__ void basis_gregory(float s, float t, Basis<D1, D2> &B)
{
    Curve cs, ct;
    bezier_curve(s, cs);
    bezier_curve(t, ct);
    constexpr int bp[12] = { 0, 1, 7, 5, 2, 6, 16, 12, 15, 17, 11, 10 };
    constexpr int bc[12] = { 0, 1, 2, 3, 0, 3, 0, 3, 0, 1, 2, 3 };
    constexpr int br[12] = { 0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 3, 3 };
    constexpr int ip[8] = { 3, 4, 8, 9, 13, 14, 18, 19 };
    constexpr int ic[8] = { 1, 1, 2, 2, 2, 2, 1, 1 };
    constexpr int ir[8] = { 1, 1, 1, 1, 2, 2, 2, 2 };
    float sc = 1.0f - s, tc = 1.0f - t;
    float den[4] = { s + t, sc + t, sc + tc, s + tc };
    float num[4] = { s, t, sc, tc };
    float g[8];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float inv = den[c] <= 0.0f ? 1.0f : 1.0f / den[c];
        g[2 * c] = num[c] * inv;
        g[2 * c + 1] = 1.0f - num[c] * inv;
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        B.p[bp[i]] = cs.b[bc[i]] * ct.b[br[i]];
        if constexpr (D1) {
            B.ds[bp[i]] = cs.d[bc[i]] * ct.b[br[i]];
            B.dt[bp[i]] = ct.d[br[i]] * cs.b[bc[i]];
        }
        if constexpr (D2) {
            B.dss[bp[i]] = cs.dd[bc[i]] * ct.b[br[i]];
            B.dst[bp[i]] = cs.d[bc[i]] * ct.d[br[i]];
            B.dtt[bp[i]] = cs.b[bc[i]] * ct.dd[br[i]];
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        B.p[ip[i]] = cs.b[ic[i]] * ct.b[ir[i]] * g[i];
        if constexpr (D1) {
            B.ds[ip[i]] = cs.d[ic[i]] * ct.b[ir[i]] * g[i];
            B.dt[ip[i]] = ct.d[ir[i]] * cs.b[ic[i]] * g[i];
        }
        if constexpr (D2) {
            B.dss[ip[i]] = cs.dd[ic[i]] * ct.b[ir[i]] * g[i];
            B.dst[ip[i]] = cs.d[ic[i]] * ct.d[ir[i]] * g[i];
            B.dtt[ip[i]] = cs.b[ic[i]] * ct.dd[ir[i]] * g[i];
        }
    }
    B.n = 20;
}

template <bool D1, bool D2>
__device__ __forceinline__ void basis_linear(float s, float t, Basis<D1, D2> &B)
{
    float u = 1.0f - s, v = 1.0f - t;
    B.p[0] = u * v; B.p[1] = s * v; B.p[2] = s * t; B.p[3] = u * t;
    if constexpr (D1) {
        B.ds[0] = -v; B.ds[1] = v; B.ds[2] = t; B.ds[3] = -t;
        B.dt[0] = -u; B.dt[1] = -s; B.dt[2] = s; B.dt[3] = u;
    }
    if constexpr (D2) {
#pragma unroll
        for (int i = 0; i < 4; ++i) { B.dss[i] = 0.0f; B.dtt[i] = 0.0f; }
        B.dst[0] = 1.0f; B.dst[1] = -1.0f; B.dst[2] = 1.0f; B.dst[3] = -1.0f;
    }
    B.n = 4;
}

}  // namespace osdhip
