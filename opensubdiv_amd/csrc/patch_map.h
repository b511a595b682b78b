// Far::PatchMap::FindPatch on the device (far/patchMap.h:193-231): one
// lane per (ptex face, u, v) query walks the face's quadtree in double
// precision, descending into the upper half of a split when the coordinate
// is >= the median, exactly as the host map does.
#pragma once
#include "../../include/osdhip.h"
#include "common.h"

namespace osdhip {

// A child slot: >= 0 a node, < 0 the leaf of patch -(c + 1), kMapEmpty unset.
constexpr int kMapEmpty = -0x7fffffff - 1;

struct PatchMapArgs {
    const v4i *nodes;        // 4 child slots per node; node f - minFace is face f's root
    const int *handles;      // 3 per patch: arrayIndex, patchIndex, vertIndex
    int minFace, maxFace, maxDepth;
    int n;
    const int *ptex;
    const double *u, *v;
    osdhip_patch_coord *coords;
};

__device__ __forceinline__ osdhip_patch_coord patch_map_lookup(const PatchMapArgs &a, int f,
                                                              double u, double v)
{
    osdhip_patch_coord c;
    c.arrayIndex = c.patchIndex = c.vertIndex = -1;
    c.s = (float)u;
    c.t = (float)v;
    if (f >= a.minFace && f <= a.maxFace && u >= 0.0 && u <= 1.0 && v >= 0.0 && v <= 1.0) {
        int node = f - a.minFace;
        double median = 0.5;
        for (int depth = 0; depth <= a.maxDepth; ++depth, median *= 0.5) {
            const int uh = u >= median, vh = v >= median;
            if (uh) u -= median;
            if (vh) v -= median;
            const v4i kids = a.nodes[node];
            // a face whose first quadrant is unset has no patches (Far
            // returns no handle for any (u, v) of it)
            if (depth == 0 && kids.x == kMapEmpty) break;
            const int q = (vh << 1) | uh;
            const int ch = q == 0 ? kids.x : q == 1 ? kids.y : q == 2 ? kids.z : kids.w;
            if (ch == kMapEmpty) break;
            if (ch < 0) {
                const int p = -(ch + 1);
                c.arrayIndex = a.handles[3 * p];
                c.patchIndex = a.handles[3 * p + 1];
                c.vertIndex = a.handles[3 * p + 2];
                break;
            }
            node = ch;
        }
    }
    return c;
}

__global__ __launch_bounds__(kBlock) void k_patch_map_find(PatchMapArgs a)
{
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= a.n) return;
    a.coords[i] = patch_map_lookup(a, a.ptex[i], a.u[i], a.v[i]);
}

// Queries given per base face in Bfr's parameterization of the face
// (bfr/parameterization.cpp:45-62, 140-175): a quad's own unit square; an
// n-gon's n quad sub-faces, sub-face k on the tile (k % uDim, k / uDim) of
// a uDim-wide grid, its normalized coordinates twice the offset into the
// tile (offsets above 0.75 belong to the next tile).  Sub-face k of base
// face f is ptex face ptexBase[f] + k, with the same normalized (s, t).
struct FaceQueryArgs {
    PatchMapArgs map;        // map.ptex unused
    const int *faces;
    const int *faceSizes, *ptexBase;
    int numFaces;
};

__device__ __forceinline__ osdhip_patch_coord face_query(const FaceQueryArgs &a, int f, double u,
                                                        double v)
{
    int ptex = -1;
    if (f >= 0 && f < a.numFaces) {
        const int n = a.faceSizes[f];
        if (n == 4) {
            ptex = a.ptexBase[f];
        } else if (n >= 3) {
            const int uDim = n < 10 ? 2 + (n > 4) : 1 + (int)sqrtf((float)(n - 1));
            int uTile = (int)u, vTile = (int)v;
            double uFrac = u - (double)uTile, vFrac = v - (double)vTile;
            if (uFrac > 0.75) { ++uTile; uFrac -= 1.0; }
            if (vFrac > 0.75) { ++vTile; vFrac -= 1.0; }
            const int k = uDim * vTile + uTile;
            if (uTile >= 0 && uTile < uDim && k >= 0 && k < n) ptex = a.ptexBase[f] + k;
            u = 2.0 * uFrac;
            v = 2.0 * vFrac;
        }
    }
    return patch_map_lookup(a.map, ptex, u, v);
}

__global__ __launch_bounds__(kBlock) void k_patch_map_find_faces(FaceQueryArgs a)
{
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= a.map.n) return;
    a.map.coords[i] = face_query(a, a.faces[i], a.map.u[i], a.map.v[i]);
}

// One (u, v) of Bfr::Tessellation's uniform pattern of a quad at `rate`
// (bfr/tessellation.cpp:158-174, 929-990): the boundary counter-clockwise
// from (0,0), 4 * rate points, then the interior rings from the outside in,
// each counter-clockwise, the innermost of an even rate being the centre.
// Along an iso-line the coordinate accumulates its step point by point in
// float, as the reference does, so the values are bit-identical.
__device__ __forceinline__ float tess_acc(float x, float d, int i)
{
    for (int j = 0; j < i; ++j) x += d;
    return x;
}

__device__ __forceinline__ void tess_quad_coord(int rate, int k, float &u, float &v)
{
    const float d = 1.0f / (float)rate;
    if (k < 4 * rate) {
        const int seg = k / rate, i = k - seg * rate;
        if (seg == 0) { u = tess_acc(0.0f, d, i); v = 0.0f; }
        else if (seg == 1) { u = 1.0f; v = tess_acc(0.0f, d, i); }
        else if (seg == 2) { u = tess_acc(1.0f, -d, i); v = 1.0f; }
        else { u = 0.0f; v = tess_acc(1.0f, -d, i); }
        return;
    }
    k -= 4 * rate;
    float u0 = d, v0 = d;
    int ur = rate - 2, vr = rate - 2;
    for (int r = 0; r < rate / 2; ++r, ur -= 2, vr -= 2, u0 += d, v0 += d) {
        if (ur > 0 && vr > 0) {
            if (k < 2 * (ur + vr)) {
                const float u1 = 1.0f - u0, v1 = 1.0f - v0;
                if (k < ur) { u = tess_acc(u0, d, k); v = v0; }
                else if ((k -= ur) < vr) { u = u1; v = tess_acc(v0, d, k); }
                else if ((k -= vr) < ur) { u = tess_acc(u1, -d, k); v = v1; }
                else { k -= ur; u = u0; v = tess_acc(v1, -d, k); }
                return;
            }
            k -= 2 * (ur + vr);
        } else if (ur > 0 || vr > 0) {
            const int n = (ur > 0 ? ur : vr) + 1;
            if (k < n) {
                if (ur > 0) { u = tess_acc(u0, d, k); v = v0; }
                else { u = u0; v = tess_acc(v0, d, k); }
                return;
            }
            k -= n;
        } else {
            u = 0.5f;
            v = 0.5f;
            return;
        }
    }
    u = v = 0.0f;   // not reached for k < (rate + 1)^2
}

struct TessArgs {
    int numFaces, rate, perFace;
    const int *ptex;
    float *uv;
    PatchMapArgs map;        // map.nodes == NULL: no PatchCoords
};

__global__ __launch_bounds__(kBlock) void k_tess_uniform_quads(TessArgs a)
{
    const long long idx = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= (long long)a.numFaces * a.perFace) return;
    const int f = (int)(idx / a.perFace), k = (int)(idx - (long long)f * a.perFace);
    float u, v;
    tess_quad_coord(a.rate, k, u, v);
    if (a.uv) {
        a.uv[2 * idx] = u;
        a.uv[2 * idx + 1] = v;
    }
    if (a.map.nodes) {
        const osdhip_patch_coord c = patch_map_lookup(a.map, a.ptex[f], (double)u, (double)v);
        a.map.coords[idx] = c;
    }
}

}  // namespace osdhip
