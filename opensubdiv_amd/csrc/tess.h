// Bfr::Tessellation coordinates (bfr/tessellation.cpp) on the device, for
// every parameterization -- quads (QUAD), Loop triangles (TRI) and the quad
// sub-faces of Catmark n-gons (QUAD_SUBFACES) -- at uniform or per-edge
// rates.
//
// The reference builds a pattern from runs of points along iso-lines: a
// run starts at (u, v) and accumulates its step in float point by point
// (appendUIsoLine / appendVIsoLine / appendUVLine, :158-185), plus single
// points (corners of sub-face rings, centres).  The host builds the run
// list of a (parameterization, rates) pattern once, computing each run's
// start and step with the reference's float expressions in the same order;
// the kernel then emits the runs of many faces, one thread per (face, run),
// accumulating sequentially, so every coordinate is the reference's float.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/osdhip.h"
#include "patch_map.h"

namespace osdhip {

struct TessRun {
    float u, v, du, dv;      // start and step; du / dv 0 when not accumulated
    int n;                   // points
    int offset;              // first point within the face's pattern
    int pad[2];
};

// A parameterization (bfr/parameterization.cpp:45-90): type 0 quad, 1
// triangle, 2 quad sub-faces on a uDim-wide grid of unit tiles.
struct TessParam {
    int type, N, uDim;
    void vertex(int k, float &u, float &v) const
    {
        if (type == 0) { u = (float)(k && k < 3); v = (float)(k > 1); }
        else if (type == 1) { u = (float)(k == 1); v = (float)(k == 2); }
        else { u = (float)(k % uDim); v = (float)(k / uDim); }
    }
};

class TessPattern {
public:
    // scheme: 0 bilinear / 1 Catmark (quads QUAD, other sizes
    // QUAD_SUBFACES), 2 Loop (triangles only, TRI).  Rates as
    // Bfr::Tessellation takes them; false when invalid.
    bool build(int scheme, int faceSize, int numRates, const int *rates)
    {
        runs.clear();
        count = 0;
        if (faceSize < 3 || numRates < 1 || !rates) return false;
        if (scheme == 2 && faceSize != 3) return false;
        if (scheme < 0 || scheme > 2) return false;
        for (int i = 0; i < numRates; ++i)
            if (rates[i] < 1) return false;
        P.N = faceSize;
        P.type = scheme == 2 ? 1 : (faceSize == 4 ? 0 : 2);
        P.uDim = faceSize < 10 ? 2 + (faceSize > 4)
                               : 1 + (int)std::sqrt((float)(faceSize - 1));
        initRates(numRates, rates);
        boundary();
        interior();
        return count == total + nInterior;
    }

    std::vector<TessRun> runs;
    int count = 0;            // coordinates per face

private:
    TessParam P;
    std::vector<int> outer;
    int inner[2] = { 0, 0 };
    bool uniform = true;
    int total = 0, nInterior = 0;

    void run(float u, float v, float du, float dv, int n)
    {
        if (n <= 0) return;
        TessRun r;
        r.u = u;
        r.v = v;
        r.du = du;
        r.dv = dv;
        r.n = n;
        r.offset = count;
        r.pad[0] = r.pad[1] = 0;
        runs.push_back(r);
        count += n;
    }
    void point(float u, float v) { run(u, v, 0.0f, 0.0f, 1); }

    // Tessellation::initializeRates (:2077-2146) and the interior point
    // counts of initializeInventoryFor* (:2165-2275)
    void initRates(int numRates, const int *rates)
    {
        const int N = P.N, cap = 32767;
        const bool isQuad = P.type == 0;
        auto c = [&](int x) { return std::min(x, cap); };
        outer.assign(N, 0);
        uniform = true;
        total = 0;
        if (numRates < N) {
            if (numRates == 2 && isQuad) {
                inner[0] = c(rates[0]);
                inner[1] = c(rates[1]);
                outer[0] = outer[2] = inner[0];
                outer[1] = outer[3] = inner[1];
                uniform = inner[0] == inner[1];
                total = 2 * (inner[0] + inner[1]);
            } else {
                inner[0] = inner[1] = c(rates[0]);
                std::fill(outer.begin(), outer.end(), inner[0]);
                total = inner[0] * N;
            }
        } else {
            for (int i = 0; i < N; ++i) {
                outer[i] = c(rates[i]);
                uniform = uniform && outer[i] == outer[0];
                total += outer[i];
            }
            if (numRates > N) {
                inner[0] = c(rates[N]);
                inner[1] = (numRates == 6 && isQuad) ? c(rates[5]) : inner[0];
                uniform = uniform && inner[0] == outer[0] && inner[1] == outer[0];
            } else if (isQuad) {
                inner[0] = (outer[0] + outer[2]) / 2;
                inner[1] = (outer[1] + outer[3]) / 2;
            } else {
                inner[0] = inner[1] = total / N;
            }
        }
        if (P.type == 0) {
            if (uniform) nInterior = inner[0] > 1 ? (inner[0] - 1) * (inner[0] - 1) : 0;
            else if (inner[0] > 1 && inner[1] > 1) nInterior = (inner[0] - 1) * (inner[1] - 1);
            else if (outer[0] == inner[0] && inner[0] == outer[2] && outer[1] == inner[1] &&
                     inner[1] == outer[3])
                nInterior = 0;                       // a segmented face
            else nInterior = 1;                      // a triangle fan around the centre
        } else if (P.type == 1) {
            const int r = inner[0];
            if (uniform) nInterior = r > 1 ? (r - 2) * (r - 1) / 2 : 0;
            else nInterior = r > 2 ? (r - 2) * (r - 1) / 2 : 1;
        } else {
            const int r = inner[0];
            if (r > 1) {
                const int e = r - 2, H = e / 2;
                nInterior = (e & 1) ? (H + 1) * (H + 1) * N + (N == 3 ? 0 : 1)
                                    : H * (H + 1) * N + 1;
            } else {
                nInterior = (uniform && N == 3) ? 0 : 1;
            }
        }
    }

    // qsub::getRingEdgeCoords (:1687-1733) with incFirst, without incLast
    void qsubEdge(int edge, int edgeRes, float t0, float dt)
    {
        const int n0 = (edgeRes - 1) / 2, n1 = (edgeRes - 1) - n0;
        float u0, v0;
        P.vertex(edge, u0, v0);
        point(u0 + t0, v0 + t0);
        if (n0) run(u0 + t0 + dt, v0 + t0, dt, 0.0f, n0);
        if (n1) {
            float u1, v1;
            P.vertex((edge + 1) % P.N, u1, v1);
            run(u1 + t0, v1 + ((edgeRes & 1) ? (0.5f - 0.5f * dt) : 0.5f), 0.0f, -dt, n1);
        }
    }

    void boundary()
    {
        if (P.type == 0) {                              // quad::GetBoundaryCoords
            run(0.0f, 0.0f, 1.0f / (float)outer[0], 0.0f, outer[0]);
            run(1.0f, 0.0f, 0.0f, 1.0f / (float)outer[1], outer[1]);
            run(1.0f, 1.0f, -1.0f / (float)outer[2], 0.0f, outer[2]);
            run(0.0f, 1.0f, 0.0f, -1.0f / (float)outer[3], outer[3]);
        } else if (P.type == 1) {                       // tri::GetBoundaryCoords
            run(0.0f, 0.0f, 1.0f / (float)outer[0], 0.0f, outer[0]);
            run(1.0f, 0.0f, -1.0f / (float)outer[1], 1.0f / (float)outer[1], outer[1]);
            run(0.0f, 1.0f, 0.0f, -1.0f / (float)outer[2], outer[2]);
        } else {                                        // qsub::GetBoundaryCoords
            for (int i = 0; i < P.N; ++i) qsubEdge(i, outer[i], 0.0f, 1.0f / (float)outer[i]);
        }
    }

    void interior()
    {
        if (nInterior == 0) return;
        if (nInterior == 1) {                           // Parameterization::GetCenterCoord
            if (P.type == 1) point(1.0f / 3.0f, 1.0f / 3.0f);
            else point(0.5f, 0.5f);
            return;
        }
        if (P.type == 0) {                              // quad::GetInteriorCoords
            const int rings = std::min(inner[0] / 2, inner[1] / 2);
            const float du = 1.0f / (float)inner[0], dv = 1.0f / (float)inner[1];
            float u = du, v = dv;
            int ur = inner[0] - 2, vr = inner[1] - 2;
            for (int i = 0; i < rings; ++i, ur -= 2, vr -= 2, u += du, v += dv) {
                if (ur > 0 && vr > 0) {
                    const float u1 = 1.0f - u, v1 = 1.0f - v;
                    run(u, v, du, 0.0f, ur);
                    run(u1, v, 0.0f, dv, vr);
                    run(u1, v1, -du, 0.0f, ur);
                    run(u, v1, 0.0f, -dv, vr);
                } else if (ur > 0) {
                    run(u, v, du, 0.0f, ur + 1);
                } else if (vr > 0) {
                    run(u, v, 0.0f, dv, vr + 1);
                } else {
                    point(0.5f, 0.5f);
                }
            }
        } else if (P.type == 1) {                       // tri::GetInteriorCoords
            const int r = inner[0], rings = r / 3;
            const float dt = 1.0f / (float)r;
            float u = dt, v = dt;
            int rr = r - 3;
            for (int i = 0; i < rings; ++i, rr -= 3, u += dt, v += dt) {
                if (rr == 0) {
                    point(1.0f / 3.0f, 1.0f / 3.0f);
                } else {
                    const float u1 = 1.0f - u * 2.0f, v1 = 1.0f - v * 2.0f;
                    run(u, v, dt, 0.0f, rr);
                    run(u1, v, -dt, dt, rr);
                    run(u, v1, 0.0f, -dt, rr);
                }
            }
        } else {                                        // qsub::GetInteriorCoords
            const int r = inner[0], rings = r / 2;
            const float dt = 1.0f / (float)r;
            float t = dt;
            int rr = r - 2;
            for (int i = 0; i < rings; ++i, rr -= 2, t += dt) {
                if (rr == 0) {
                    point(0.5f, 0.5f);
                } else if (rr == 1) {
                    for (int k = 0; k < P.N; ++k) {
                        float u0, v0;
                        P.vertex(k, u0, v0);
                        point(u0 + t, v0 + t);
                    }
                    if (P.N != 3) point(0.5f, 0.5f);
                } else {
                    for (int k = 0; k < P.N; ++k) qsubEdge(k, rr, t, dt);
                }
            }
        }
    }
};

struct TessRunArgs {
    const TessRun *runs;
    int numRuns, perFace, numFaces;
    float *uv;                       // face-major (u, v) pairs, or NULL
    // PatchCoords through Bfr's face parameterization (map.nodes NULL: none)
    FaceQueryArgs q;                 // q.faces: the base face of each pattern
};

// One thread per (face, run): the run's points accumulate their step in
// float, point by point, as the reference's iso-lines do.
__global__ __launch_bounds__(kBlock) void k_tess_runs(TessRunArgs a)
{
    const long long idx = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= (long long)a.numFaces * a.numRuns) return;
    const int f = (int)(idx / a.numRuns), r = (int)(idx - (long long)f * a.numRuns);
    const TessRun run = a.runs[r];
    const long long base = (long long)f * a.perFace + run.offset;
    const bool accU = run.du != 0.0f, accV = run.dv != 0.0f;
    float u = run.u, v = run.v;
    for (int i = 0; i < run.n; ++i) {
        if (a.uv) {
            a.uv[2 * (base + i)] = u;
            a.uv[2 * (base + i) + 1] = v;
        }
        if (a.q.map.nodes)
            a.q.map.coords[base + i] = face_query(a.q, a.q.faces[f], (double)u, (double)v);
        if (accU) u += run.du;
        if (accV) v += run.dv;
    }
}

}  // namespace osdhip
