// Host-side patch basis (double or float) for limit stencils: the same
// bases the device kernel evaluates (osd/patchBasis.h:55-490, 1447-1610),
// REGULAR bicubic B-spline with the boundary fold, GREGORY_BASIS with the
// approximate derivatives, QUADS bilinear.  u, v are ptex coordinates;
// derivatives are returned in ptex space (x 2^depth, x 4^depth).
#pragma once
#include <cstdint>

namespace osdfar {

template <typename R>
inline void bsplineCurve(R t, R b[4], R d[4], R dd[4])
{
    const R k6 = R(1) / R(6);
    R t2 = t * t, t3 = t * t2;
    b[0] = k6 * (R(1) - R(3) * (t - t2) - t3);
    b[1] = k6 * (R(4) - R(6) * t2 + R(3) * t3);
    b[2] = k6 * (R(1) + R(3) * (t + t2 - t3));
    b[3] = k6 * t3;
    d[0] = -R(0.5) * t2 + t - R(0.5);
    d[1] = R(1.5) * t2 - R(2) * t;
    d[2] = -R(1.5) * t2 + t + R(0.5);
    d[3] = R(0.5) * t2;
    dd[0] = -t + R(1);
    dd[1] = R(3) * t - R(2);
    dd[2] = -R(3) * t + R(1);
    dd[3] = t;
}

template <typename R>
inline void bezierCurve(R t, R b[4], R d[4], R dd[4])
{
    R t2 = t * t, c = R(1) - t, c2 = c * c;
    b[0] = c2 * c;
    b[1] = c2 * t * R(3);
    b[2] = t2 * c * R(3);
    b[3] = t2 * t;
    d[0] = -R(3) * c2;
    d[1] = R(9) * t2 - R(12) * t + R(3);
    d[2] = -R(9) * t2 + R(6) * t;
    d[3] = R(3) * t2;
    dd[0] = R(6) * c;
    dd[1] = R(18) * t - R(12);
    dd[2] = -R(18) * t + R(6);
    dd[3] = R(6) * t;
}

template <typename R>
inline void foldBoundary(int mask, R *w)
{
    if (!w) return;
    if (mask & 1) for (int c = 0; c < 4; ++c) { w[8 + c] -= w[c]; w[4 + c] += R(2) * w[c]; w[c] = 0; }
    if (mask & 2) for (int r = 0; r < 16; r += 4) { w[r + 1] -= w[r + 3]; w[r + 2] += R(2) * w[r + 3]; w[r + 3] = 0; }
    if (mask & 4) for (int c = 0; c < 4; ++c) { w[4 + c] -= w[12 + c]; w[8 + c] += R(2) * w[12 + c]; w[12 + c] = 0; }
    if (mask & 8) for (int r = 0; r < 16; r += 4) { w[r + 2] -= w[r]; w[r + 1] += R(2) * w[r]; w[r] = 0; }
}

// Returns the number of points (16, 20, 4) or 0 for another type.
template <typename R>
inline int patchBasis(int type, uint32_t field1, R u, R v, R *wP, R *wDs, R *wDt,
                      R *wDss, R *wDst, R *wDtt)
{
    const int depth = (int)(field1 & 0xf), nonquad = (int)((field1 >> 4) & 1);
    const int boundary = (int)((field1 >> 7) & 0x1f);
    const int pv = (int)((field1 >> 12) & 0x3ff), pu = (int)((field1 >> 22) & 0x3ff);
    const R scale = R(1 << (depth - nonquad));
    const R s = u * scale - R(pu), t = v * scale - R(pv);
    const bool d1 = wDs && wDt, d2 = d1 && wDss && wDst && wDtt;
    R bs[4], ds[4], dss[4], bt[4], dt[4], dtt[4];
    int n = 0;
    if (type == 6) {
        bsplineCurve(s, bs, ds, dss);
        bsplineCurve(t, bt, dt, dtt);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                const int k = 4 * i + j;
                wP[k] = bs[j] * bt[i];
                if (d1) { wDs[k] = ds[j] * bt[i]; wDt[k] = bs[j] * dt[i]; }
                if (d2) { wDss[k] = dss[j] * bt[i]; wDst[k] = ds[j] * dt[i]; wDtt[k] = bs[j] * dtt[i]; }
            }
        if (boundary) {
            foldBoundary(boundary, wP);
            if (d1) { foldBoundary(boundary, wDs); foldBoundary(boundary, wDt); }
            if (d2) { foldBoundary(boundary, wDss); foldBoundary(boundary, wDst); foldBoundary(boundary, wDtt); }
        }
        n = 16;
    } else if (type == 9) {
        static const int bp[12] = { 0, 1, 7, 5, 2, 6, 16, 12, 15, 17, 11, 10 };
        static const int bc[12] = { 0, 1, 2, 3, 0, 3, 0, 3, 0, 1, 2, 3 };
        static const int br[12] = { 0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 3, 3 };
        static const int ip[8] = { 3, 4, 8, 9, 13, 14, 18, 19 };
        static const int ic[8] = { 1, 1, 2, 2, 2, 2, 1, 1 };
        static const int ir[8] = { 1, 1, 1, 1, 2, 2, 2, 2 };
        bezierCurve(s, bs, ds, dss);
        bezierCurve(t, bt, dt, dtt);
        const R sc = R(1) - s, tc = R(1) - t;
        const R den[4] = { s + t, sc + t, sc + tc, s + tc }, num[4] = { s, t, sc, tc };
        R g[8];
        for (int c = 0; c < 4; ++c) {
            const R inv = den[c] <= R(0) ? R(1) : R(1) / den[c];
            g[2 * c] = num[c] * inv;
            g[2 * c + 1] = R(1) - num[c] * inv;
        }
        for (int i = 0; i < 12; ++i) {
            const int p = bp[i], c = bc[i], r = br[i];
            wP[p] = bs[c] * bt[r];
            if (d1) { wDs[p] = ds[c] * bt[r]; wDt[p] = dt[r] * bs[c]; }
            if (d2) { wDss[p] = dss[c] * bt[r]; wDst[p] = ds[c] * dt[r]; wDtt[p] = bs[c] * dtt[r]; }
        }
        for (int i = 0; i < 8; ++i) {
            const int p = ip[i], c = ic[i], r = ir[i];
            wP[p] = bs[c] * bt[r] * g[i];
            if (d1) { wDs[p] = ds[c] * bt[r] * g[i]; wDt[p] = dt[r] * bs[c] * g[i]; }
            if (d2) {
                wDss[p] = dss[c] * bt[r] * g[i];
                wDst[p] = ds[c] * dt[r] * g[i];
                wDtt[p] = bs[c] * dtt[r] * g[i];
            }
        }
        n = 20;
    } else if (type == 3) {
        const R a = R(1) - s, b = R(1) - t;
        wP[0] = a * b; wP[1] = s * b; wP[2] = s * t; wP[3] = a * t;
        if (d1) {
            wDs[0] = -b; wDs[1] = b; wDs[2] = t; wDs[3] = -t;
            wDt[0] = -a; wDt[1] = -s; wDt[2] = s; wDt[3] = a;
        }
        if (d2)
            for (int i = 0; i < 4; ++i) { wDss[i] = 0; wDtt[i] = 0; wDst[i] = (i & 1) ? R(-1) : R(1); }
        n = 4;
    } else {
        return 0;
    }
    if (d1) {
        const R k1 = R(1 << depth);
        for (int i = 0; i < n; ++i) { wDs[i] *= k1; wDt[i] *= k1; }
        if (d2) {
            const R k2 = k1 * k1;
            for (int i = 0; i < n; ++i) { wDss[i] *= k2; wDst[i] *= k2; wDtt[i] *= k2; }
        }
    }
    return n;
}

}  // namespace osdfar
