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

__global__ __launch_bounds__(kBlock) void k_patch_map_find(PatchMapArgs a)
{
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= a.n) return;
    const int f = a.ptex[i];
    double u = a.u[i], v = a.v[i];
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
    a.coords[i] = c;
}

}  // namespace osdhip
