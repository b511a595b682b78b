// Device-side patch basis for the limit-evaluation kernel (gfx950).
//
// Evaluates, in registers, the weights that OpenSubdiv's
// OsdEvaluatePatchBasis (osd/patchBasis.h:1447-1610) produces for every
// patch type it evaluates: REGULAR bicubic B-spline with the boundary fold
// (patchBasis.h:141-245), GREGORY_BASIS 20-point Gregory with the default
// approximate derivatives (patchBasis.h:334-490), QUADS bilinear
// (patchBasis.h:55-97), and the triangular types of Loop meshes: LOOP
// quartic box spline with its boundary phantom points (patchBasis.h:
// 527-850), GREGORY_TRIANGLE 18-point Gregory (patchBasis.h:852-1170) and
// TRIANGLES linear (patchBasis.h:494-524).  Weights are produced UNSCALED;
// the caller applies the +-2^depth / +-4^depth derivative factors to the
// accumulated sums, which is exact because the factors are signed powers
// of two.
//
// Parity: the translation unit is compiled with -ffp-contract=off and every
// product / sum is written in the reference's order, so results match
// Osd::CpuEvaluator bit for bit.
//
// All arrays are fixed-size and indexed by compile-time constants after
// unrolling, so they live in VGPRs.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/osdhip.h"

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


// ---- triangular patch types ------------------------------------------------

template <bool D1, bool D2>
__device__ __forceinline__ void basis_linear_tri(float s, float t, Basis<D1, D2> &B)
{
    B.p[0] = 1.0f - s - t; B.p[1] = s; B.p[2] = t;
    if constexpr (D1) {
        B.ds[0] = -1.0f; B.ds[1] = 1.0f; B.ds[2] = 0.0f;
        B.dt[0] = -1.0f; B.dt[1] = 0.0f; B.dt[2] = 1.0f;
    }
    if constexpr (D2) {
#pragma unroll
        for (int i = 0; i < 3; ++i) { B.dss[i] = 0.0f; B.dst[i] = 0.0f; B.dtt[i] = 0.0f; }
    }
    B.n = 3;
}

// One derivative of the 12 box-spline weights: weight i is
// S * sum_k c[d][i][k] * M[k] over the quartic monomials M, summed in
// ascending k from the first nonzero term (the reference's term order), for
// d = 0..5 <-> s^a t^b derivatives 00, 10, 01, 20, 11, 02.
__device__ __forceinline__ void box_weights(const float (&M)[15], int d, float (&w)[20])
{
    constexpr signed char c[6][12][15] = {
        {{1, -2, -4, 0, 6, 6, 2, 0, -6, -4, -1, -2, 0, 2, 1},
         {1, 2, -2, 0, -6, 0, -4, 0, 6, 2, 2, 4, 0, -2, -1},
         {0, 0, 0, 0, 0, 0, 2, 0, 0, 0, -1, -2, 0, 0, 0},
         {1, -4, -2, 6, 6, 0, -4, -6, 0, 2, 1, 2, 0, -2, -1},
         {6, 0, 0, -12, -12, -12, 8, 12, 12, 8, -1, -2, 0, -2, -1},
         {1, 4, 2, 6, 6, 0, -4, -6, -12, -4, -1, -2, 0, 4, 2},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 0, 0, 0},
         {1, -2, 2, 0, -6, 0, 2, 6, 0, -4, -1, -2, 0, 4, 2},
         {1, 2, 4, 0, 6, 6, -4, -12, -6, -4, 2, 4, 0, -2, -1},
         {0, 0, 0, 0, 0, 0, 2, 6, 6, 2, -1, -2, 0, -2, -1},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 0, 0, 0, -2, -1},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2, 1}},
        {{-1, 0, 3, 3, 0, -3, -2, -3, 0, 1, 0, 0, 0, 0, 0},
         {1, 0, -3, -6, 0, 3, 4, 6, 0, -1, 0, 0, 0, 0, 0},
         {0, 0, 0, 3, 0, 0, -2, -3, 0, 0, 0, 0, 0, 0, 0},
         {-2, 6, 3, -6, -6, 0, 2, 3, 0, -1, 0, 0, 0, 0, 0},
         {0, -12, -6, 12, 12, 6, -2, -3, 0, -1, 0, 0, 0, 0, 0},
         {2, 6, 3, -6, -6, -6, -2, -3, 0, 2, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 2, 3, 0, 0, 0, 0, 0, 0, 0},
         {-1, 0, -3, 3, 6, 0, -2, -3, 0, 2, 0, 0, 0, 0, 0},
         {1, 0, 3, -6, -12, -3, 4, 6, 0, -1, 0, 0, 0, 0, 0},
         {0, 0, 0, 3, 6, 3, -2, -3, 0, -1, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0}},
        {{-2, 3, 6, 0, -6, -6, -1, 0, 3, 2, 0, 0, 0, 0, 0},
         {-1, -3, 0, 0, 6, 3, 2, 0, -3, -2, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0},
         {-1, 3, 0, -3, 0, 3, 1, 0, -3, -2, 0, 0, 0, 0, 0},
         {0, -6, -12, 6, 12, 12, -1, 0, -3, -2, 0, 0, 0, 0, 0},
         {1, 3, 0, -3, -12, -6, -1, 0, 6, 4, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0},
         {1, -3, 0, 3, 0, -6, -1, 0, 6, 4, 0, 0, 0, 0, 0},
         {2, 3, 6, -6, -6, -6, 2, 0, -3, -2, 0, 0, 0, 0, 0},
         {0, 0, 0, 3, 6, 3, -1, 0, -3, -2, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 3, 0, 0, -3, -2, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 0, 0, 3, 2, 0, 0, 0, 0, 0}},
        {{0, 1, 0, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, -2, 0, 2, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 1, 0, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {1, -2, -1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {-2, 4, 2, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {1, -2, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, -2, -2, 2, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}},
        {{1, 0, -2, -1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {-1, 0, 2, 2, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {1, -2, 0, 1, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {-2, 4, 4, -1, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {1, -2, -4, -1, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {-1, 2, 0, -1, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {1, -4, -2, 2, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 2, 2, -1, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0}},
        {{1, -1, -2, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 1, 1, 0, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 1, 0, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {-2, 2, 4, 0, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, -2, -2, 0, 2, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, -2, 0, 2, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {1, -1, -2, 0, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 1, 1, 0, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 1, 0, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0},
         {0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0}}};
    const float S = d == 0 ? 1.0f * (1.0f / 12.0f) : d < 3 ? 1.0f * (1.0f / 6.0f)
                  : d == 4 ? 1.0f * (1.0f / 2.0f) : 1.0f;
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        float acc = 0.0f;
        bool any = false;
#pragma unroll
        for (int k = 0; k < 15; ++k) {
            const int ck = c[d][i][k];
            if (ck == 0) continue;
            const float term = k == 0 ? (float)ck : (float)ck * M[k];
            acc = any ? acc + term : term;
            any = true;
        }
        w[i] = any ? S * acc : 0.0f;
    }
}

// The box spline's boundary phantom points: a boundary edge's three
// phantom points and a boundary vertex's two move their weights onto the
// patch's points (each as P = B + (B' - I)), edges 0, 1, 2 then vertices
// 0, 1, 2, every step reading the weights the previous one left.  Edge or
// vertex e is edge or vertex 0 seen through the corner rotation rot^e (point
// i seen from corner 0 is point rot[i] seen from corner 1).  The point
// indices are compile-time constants after unrolling.
__device__ __forceinline__ void box_boundary(int mask, float (&w)[20])
{
    constexpr int rot[12] = { 6, 9, 11, 2, 5, 8, 10, 1, 4, 7, 0, 3 };
    const int upper = (mask >> 3) & 3;
    int ebits = mask & 7, vbits = 0;
    if (upper == 1) { vbits = ebits; ebits = 0; }
    else if (upper == 2) vbits = ((ebits & 1) << 2) | (ebits >> 1);
    int p[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) p[i] = i;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        if (ebits & (1 << e)) {
            const bool prev = (ebits >> ((e + 2) % 3)) & 1, next = (ebits >> ((e + 1) % 3)) & 1;
            const float a = w[p[0]];
            if (prev) { w[p[4]] += a; w[p[4]] += a; w[p[8]] -= a; }
            else { w[p[4]] += a; w[p[3]] += a; w[p[7]] -= a; }
            const float b = w[p[1]];
            w[p[4]] += b; w[p[5]] += b; w[p[8]] -= b;
            const float c = w[p[2]];
            if (next) { w[p[5]] += c; w[p[5]] += c; w[p[8]] -= c; }
            else { w[p[5]] += c; w[p[6]] += c; w[p[9]] -= c; }
            w[p[0]] = 0.0f; w[p[1]] = 0.0f; w[p[2]] = 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 12; ++i) p[i] = rot[p[i]];
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        if (vbits & (1 << e)) {
            const float a = w[p[3]];
            w[p[4]] += a; w[p[7]] += a; w[p[8]] -= a;
            const float b = w[p[0]];
            w[p[4]] += b; w[p[1]] += b; w[p[5]] -= b;
            w[p[3]] = 0.0f; w[p[0]] = 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 12; ++i) p[i] = rot[p[i]];
    }
}

template <bool D1, bool D2>
__device__ __forceinline__ void basis_box_tri(float s, float t, int boundary, Basis<D1, D2> &B)
{
    float M[15];
    M[0] = 1.0f; M[1] = s; M[2] = t;
    M[3] = s * s; M[4] = s * t; M[5] = t * t;
    M[6] = M[3] * s; M[7] = M[4] * s; M[8] = M[4] * t; M[9] = M[5] * t;
    M[10] = M[6] * s; M[11] = M[7] * s; M[12] = M[3] * M[5]; M[13] = M[8] * t; M[14] = M[9] * t;
    box_weights(M, 0, B.p);
    if constexpr (D1) { box_weights(M, 1, B.ds); box_weights(M, 2, B.dt); }
    if constexpr (D2) {
        box_weights(M, 3, B.dss);
        box_weights(M, 4, B.dst);
        box_weights(M, 5, B.dtt);
    }
    if (boundary) {
        box_boundary(boundary, B.p);
        if constexpr (D1) { box_boundary(boundary, B.ds); box_boundary(boundary, B.dt); }
        if constexpr (D2) {
            box_boundary(boundary, B.dss);
            box_boundary(boundary, B.dst);
            box_boundary(boundary, B.dtt);
        }
    }
    B.n = 12;
}

// Quartic Bezier triangle weights (15 points) in barycentric u = s, v = t,
// w = 1 - u - v: d = 0 value, 1 d/ds, 2 d/dt, 3 d2/ds2, 4 d2/dsdt,
// 5 d2/dt2, in the forms and operation order of osd/patchBasis.h:852-1020.
__device__ __forceinline__ void bezier_tri(float s, float t, int d, float (&b)[15])
{
    const float u = s, v = t, w = 1.0f - u - v;
    const float uu = u * u, vv = v * v, ww = w * w, uv = u * v, vw = v * w, uw = u * w;
    if (d == 0) {
        b[0] = ww * ww;            b[1] = 4.0f * uw * ww;     b[2] = 6.0f * uw * uw;
        b[3] = 4.0f * uw * uu;     b[4] = uu * uu;            b[5] = 4.0f * vw * ww;
        b[6] = 12.0f * ww * uv;    b[7] = 12.0f * uu * vw;    b[8] = 4.0f * uv * uu;
        b[9] = 6.0f * vw * vw;     b[10] = 12.0f * vv * uw;   b[11] = 6.0f * uv * uv;
        b[12] = 4.0f * vw * vv;    b[13] = 4.0f * uv * vv;    b[14] = vv * vv;
    } else if (d == 1) {
        b[0] = -4.0f * ww * w;                 b[1] = 4.0f * ww * (w - 3.0f * u);
        b[2] = 12.0f * uw * (w - u);           b[3] = 4.0f * uu * (3.0f * w - u);
        b[4] = 4.0f * uu * u;                  b[5] = -12.0f * vw * w;
        b[6] = 12.0f * vw * (w - 2.0f * u);    b[7] = 12.0f * uv * (2.0f * w - u);
        b[8] = 12.0f * uv * u;                 b[9] = -12.0f * vv * w;
        b[10] = 12.0f * vv * (w - u);          b[11] = 12.0f * vv * u;
        b[12] = -4.0f * vv * v;                b[13] = 4.0f * vv * v;
        b[14] = 0.0f;
    } else if (d == 2) {
        b[0] = -4.0f * ww * w;                 b[1] = -12.0f * ww * u;
        b[2] = -12.0f * uu * w;                b[3] = -4.0f * uu * u;
        b[4] = 0.0f;                           b[5] = 4.0f * ww * (w - 3.0f * v);
        b[6] = 12.0f * uw * (w - 2.0f * v);    b[7] = 12.0f * uu * (w - v);
        b[8] = 4.0f * uu * u;                  b[9] = 12.0f * vw * (w - v);
        b[10] = 12.0f * uv * (2.0f * w - v);   b[11] = 12.0f * uv * u;
        b[12] = 4.0f * vv * (3.0f * w - v);    b[13] = 12.0f * vv * u;
        b[14] = 4.0f * vv * v;
    } else if (d == 3) {
        b[0] = 12.0f * ww;                     b[1] = 24.0f * (uw - ww);
        b[2] = 12.0f * (uu - 4.0f * uw + ww);  b[3] = 24.0f * (uw - uu);
        b[4] = 12.0f * uu;                     b[5] = 24.0f * vw;
        b[6] = 24.0f * (uv - 2.0f * vw);       b[7] = 24.0f * (vw - 2.0f * uv);
        b[8] = 24.0f * uv;                     b[9] = 12.0f * vv;
        b[10] = -24.0f * vv;                   b[11] = 12.0f * vv;
        b[12] = 0.0f; b[13] = 0.0f; b[14] = 0.0f;
    } else if (d == 5) {
        b[0] = 12.0f * ww;                     b[1] = 24.0f * uw;
        b[2] = 12.0f * uu;                     b[3] = 0.0f; b[4] = 0.0f;
        b[5] = 24.0f * (vw - ww);              b[6] = 24.0f * (uv - 2.0f * uw);
        b[7] = -24.0f * uu;                    b[8] = 0.0f;
        b[9] = 12.0f * (vv - 4.0f * vw + ww);  b[10] = 24.0f * (uw - 2.0f * uv);
        b[11] = 12.0f * uu;                    b[12] = 24.0f * (vw - vv);
        b[13] = 24.0f * uv;                    b[14] = 12.0f * vv;
    } else {   // the mixed derivative, partly from its own entries
        b[0] = 12.0f * ww;
        b[3] = -12.0f * uu;
        b[13] = 12.0f * vv;
        b[11] = 24.0f * uv;
        b[1] = 24.0f * uw - b[0];
        b[2] = -24.0f * uw - b[3];
        b[5] = 24.0f * vw - b[0];
        b[6] = -24.0f * vw + b[11] - b[1];
        b[8] = -b[3];
        b[7] = -(b[11] + b[2]);
        b[9] = b[13] - b[5] - b[0];
        b[10] = -(b[9] + b[11]);
        b[12] = -b[13];
        b[4] = 0.0f; b[14] = 0.0f;
    }
}

// Gregory point k takes Bezier weight gb[k], times the rational factor
// g[gg[k]] for the six face points (pairs splitting the three interior
// Bezier points; each pair sums to one, 1 / 0 where its denominator
// vanishes).  Derivatives hold g fixed: the approximate derivatives.
__device__ __forceinline__ void bezier_to_gregory_tri(const float (&b)[15], const float (&g)[6],
                                                      float (&w)[20])
{
    constexpr int gb[18] = { 0, 1, 5, 6, 6, 4, 8, 3, 7, 7, 14, 12, 13, 10, 10, 2, 11, 9 };
    constexpr int gg[18] = { -1, -1, -1, 0, 1, -1, -1, -1, 2, 3, -1, -1, -1, 4, 5, -1, -1, -1 };
#pragma unroll
    for (int k = 0; k < 18; ++k) w[k] = gg[k] < 0 ? b[gb[k]] : b[gb[k]] * g[gg[k]];
}

template <bool D1, bool D2>
__device__ __forceinline__ void basis_gregory_tri(float s, float t, Basis<D1, D2> &B)
{
    const float u = s, v = t, w = 1.0f - u - v;
    float g[6] = { 1.0f, 0.0f, 1.0f, 0.0f, 1.0f, 0.0f };
    if (u + v > 0.0f) { g[0] = u / (u + v); g[1] = v / (u + v); }
    if (v + w > 0.0f) { g[2] = v / (v + w); g[3] = w / (v + w); }
    if (w + u > 0.0f) { g[4] = w / (w + u); g[5] = u / (w + u); }
    float b[15];
    bezier_tri(s, t, 0, b);
    bezier_to_gregory_tri(b, g, B.p);
    if constexpr (D1) {
        bezier_tri(s, t, 1, b);
        bezier_to_gregory_tri(b, g, B.ds);
        bezier_tri(s, t, 2, b);
        bezier_to_gregory_tri(b, g, B.dt);
    }
    if constexpr (D2) {
        bezier_tri(s, t, 3, b);
        bezier_to_gregory_tri(b, g, B.dss);
        bezier_tri(s, t, 4, b);
        bezier_to_gregory_tri(b, g, B.dst);
        bezier_tri(s, t, 5, b);
        bezier_to_gregory_tri(b, g, B.dtt);
    }
    B.n = 18;
}

// ---- PatchParam decode and the per-type point count ----------------------

// Points of a patch type (REGULAR 16, GREGORY_BASIS 20, QUADS 4, LOOP 12,
// GREGORY_TRIANGLE 18, TRIANGLES 3); 0 for a type the reference does not
// evaluate (its outputs are then zero, as CpuEvaluator's Clear() leaves
// them).
__device__ __forceinline__ int patch_points(int type)
{
    return type == OSDHIP_PATCH_REGULAR ? 16
         : type == OSDHIP_PATCH_GREGORY_BASIS ? 20
         : type == OSDHIP_PATCH_QUADS ? 4
         : type == OSDHIP_PATCH_LOOP ? 12
         : type == OSDHIP_PATCH_GREGORY_TRIANGLE ? 18
         : type == OSDHIP_PATCH_TRIANGLES ? 3 : 0;
}

// PatchParam Normalize (osd/patchBasisTypes.h:389-397) of ptex-face (s, t)
// to patch-local coordinates, or for the triangular types NormalizeTriangle
// (:413-426): a rotated triangle (u + v >= 2^depth) is mirrored and its
// derivatives change sign.  Returns the first-derivative factor
// +-2^depth; the second-derivative factor is sign * factor^2.
__device__ __forceinline__ float patch_normalize(int type, unsigned field1, float &s, float &t,
                                                 float &sign)
{
    const int depth = (int)(field1 & 0xfu);
    const int nonquad = (int)((field1 >> 4) & 1u);
    const int pv = (int)((field1 >> 12) & 0x3ffu);
    const int pu = (int)((field1 >> 22) & 0x3ffu);
    const float frac_inv = (float)(1 << (depth - nonquad));
    const bool tri = type == OSDHIP_PATCH_LOOP || type == OSDHIP_PATCH_GREGORY_TRIANGLE ||
                     type == OSDHIP_PATCH_TRIANGLES;
    sign = 1.0f;
    if (tri && pu + pv >= (1 << depth)) {
        s = (float)((1 << depth) - pu) - s * frac_inv;
        t = (float)((1 << depth) - pv) - t * frac_inv;
        sign = -1.0f;
    } else {
        s = s * frac_inv - (float)pu;
        t = t * frac_inv - (float)pv;
    }
    return sign * (float)(1 << depth);
}

}  // namespace osdhip
