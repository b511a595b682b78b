// Batched evaluation: many meshes, each with many frames of control points,
// in one graph launch of two kernels.
//
//   phase A  k_batch_stencils  every stencil segment (one record-layout
//                              table applied to F frames): refinement
//                              tables and the patch-point tables that fill
//                              each mesh's vertex buffer [control | points];
//   phase B  k_batch_patches   every patch segment (one mesh's PatchCoords
//                              evaluated on its F frames), reading the
//                              vertex buffers phase A completed.
//
// Frames are extra right-hand-side columns: a wave loads a table slice's
// records once per chunk of FC frames (later chunks hit L1/L2), and a
// PatchCoord's basis is evaluated once for all frames.  Element (v, f, k)
// of a multi-frame buffer lives at base + v * vStride + f * fStride + k, so
// both frame-major and vertex-major (frames interleaved) layouts work.
//
// Work is cut into items of up to `itemSlices` 64-row slices (phase A) or
// kBlock coords (phase B).  Static scheduling maps one item to one block;
// persistent scheduling launches a fixed grid whose blocks claim items from
// a device counter (one vector atomic per item, by lane 0 of wave 0) until
// the list is exhausted -- every block exits when the counter passes the
// item count.
//
// Every sum accumulates w * x in the same order as the single-mesh kernels
// (k_stencils_rec / k_patches), so batched results equal one-mesh-per-launch
// results bit for bit.
#pragma once
#include "../../include/osdhip.h"
#include "common.h"
#include "patch_basis_device.h"
#include "stencil_sliced.h"

namespace osdhip {

struct BatchStencilSeg {
    const unsigned *rec;          // record layout, one weight array (R = 2)
    const long long *sliceOff;
    const int *sliceWidth;
    const int *rowSize;           // padded to whole slices
    const float *src;
    long long srcV, srcF;
    float *dst;
    long long dstV, dstF;
    int nrows, nslices, frames, pad;
};

struct BatchPatchSeg {
    const osdhip_patch_coord *coords;
    const osdhip_patch_array *arrays;
    const int *indices;
    const osdhip_patch_param *params;
    const float *src;
    long long srcV, srcF;
    float *out[3];                // P, dPdu, dPdv (NULL = skip)
    long long outV, outF;
    int ncoords, frames;
};

struct BatchItem {
    int seg, first, count, pad;   // phase A: slices [first, first+count); B: coords
};

struct BatchArgs {
    const BatchStencilSeg *sseg;
    const BatchPatchSeg *pseg;
    const BatchItem *items;
    int nitems;
    int persistent;
    unsigned *counter;            // persistent: next item
};

__global__ void k_batch_reset(unsigned *counters, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) counters[i] = 0u;
}

// Next item of this block: static -> blockIdx.x once; persistent -> claimed
// from the counter.  Returns -1 when done.  Must be reached by every thread
// of the block (it contains __syncthreads).
__device__ __forceinline__ int batch_next_item(const BatchArgs &a, int iter)
{
    if (!a.persistent) return iter == 0 && (int)blockIdx.x < a.nitems ? (int)blockIdx.x : -1;
    __shared__ int slot;
    __syncthreads();                          // previous item's readers are done
    if (threadIdx.x == 0) slot = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    const int it = slot;
    return it < a.nitems ? it : -1;
}

template <int LC, int FC, int G>
__global__ __launch_bounds__(kBlock) void k_batch_stencils(BatchArgs a)
{
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wv = (int)threadIdx.x / kWave;
    for (int iter = 0;; ++iter) {
        const int itx = batch_next_item(a, iter);
        if (itx < 0) break;
        const BatchItem item = a.items[itx];
        const BatchStencilSeg sg = a.sseg[item.seg];
        for (int sl = wv; sl < item.count; sl += kWavesPerBlock) {
            const int s = item.first + sl;
            const int row = s * kWave + lane;
            const bool active = row < sg.nrows;
            const int sz = sg.rowSize[row];
            const int W = sg.sliceWidth[s];
            const unsigned *rp = sg.rec + (size_t)(sg.sliceOff[s] + lane) * 2;
            for (int f0 = 0; f0 < sg.frames; f0 += FC) {
                float acc[FC][LC];
#pragma unroll
                for (int c = 0; c < FC; ++c)
#pragma unroll
                    for (int k = 0; k < LC; ++k) acc[c][k] = 0.0f;
                for (int j = 0; j < W; j += G) {
                    int ii[G];
                    float ww[1][G];
#pragma unroll
                    for (int u = 0; u < G; ++u) {
                        ii[u] = 0;
                        ww[0][u] = 0.0f;
                        if (j + u < W)                          // wave-uniform
                            load_record<1, G>(rp + (size_t)(j + u) * (kWave * 2), ii[u], ww, u);
                    }
                    float v[G][FC][LC];
#pragma unroll
                    for (int u = 0; u < G; ++u) {
                        const float *p = sg.src + (size_t)(unsigned)ii[u] * sg.srcV +
                                         (size_t)f0 * sg.srcF;
#pragma unroll
                        for (int c = 0; c < FC; ++c) {
                            if (f0 + c < sg.frames) {           // wave-uniform
#pragma unroll
                                for (int k = 0; k < LC; ++k) v[u][c][k] = p[c * sg.srcF + k];
                            } else {
#pragma unroll
                                for (int k = 0; k < LC; ++k) v[u][c][k] = 0.0f;
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < G; ++u) {
                        const bool take = j + u < sz;
#pragma unroll
                        for (int c = 0; c < FC; ++c)
#pragma unroll
                            for (int k = 0; k < LC; ++k) {
                                const float t = acc[c][k] + v[u][c][k] * ww[0][u];
                                acc[c][k] = take ? t : acc[c][k];
                            }
                    }
                }
                if (active) {
#pragma unroll
                    for (int c = 0; c < FC; ++c) {
                        if (f0 + c >= sg.frames) break;
                        float *d = sg.dst + (size_t)row * sg.dstV + (size_t)(f0 + c) * sg.dstF;
#pragma unroll
                        for (int k = 0; k < LC; ++k) d[k] = acc[c][k];
                    }
                }
            }
        }
    }
}

template <int LC, int FC>
__global__ __launch_bounds__(kBlock) void k_batch_patches(BatchArgs a)
{
    for (int iter = 0;; ++iter) {
        const int itx = batch_next_item(a, iter);
        if (itx < 0) break;
        const BatchItem item = a.items[itx];
        const BatchPatchSeg sg = a.pseg[item.seg];
        const int i = item.first + (int)threadIdx.x;
        if (i >= item.first + item.count) continue;

        const osdhip_patch_coord c = sg.coords[i];
        const osdhip_patch_array arr = sg.arrays[c.arrayIndex];
        const osdhip_patch_param pp = sg.params[c.patchIndex];
        const unsigned f1 = pp.field1;
        const int type = ((f1 >> 5) & 1u) ? arr.regDesc : arr.desc;
        const int depth = (int)(f1 & 0xfu);
        const int nonquad = (int)((f1 >> 4) & 1u);
        const int boundary = (int)((f1 >> 7) & 0x1fu);
        const int pv = (int)((f1 >> 12) & 0x3ffu);
        const int pu = (int)((f1 >> 22) & 0x3ffu);
        const float frac_inv = (float)(1 << (depth - nonquad));
        const float s = c.s * frac_inv - (float)pu;
        const float t = c.t * frac_inv - (float)pv;
        const int n = type == OSDHIP_PATCH_REGULAR ? 16
                    : type == OSDHIP_PATCH_GREGORY_BASIS ? 20
                    : type == OSDHIP_PATCH_QUADS ? 4 : 0;
        const int *cvs = sg.indices + arr.indexBase + arr.stride * (c.patchIndex - arr.primitiveIdBase);
        Basis<true, false> B;
        if (n == 16) basis_bspline<true, false>(s, t, boundary, B);
        else if (n == 20) basis_gregory<true, false>(s, t, B);
        else if (n == 4) basis_linear<true, false>(s, t, B);
        else {
#pragma unroll
            for (int j = 0; j < 20; ++j) B.p[j] = B.ds[j] = B.dt[j] = 0.0f;
        }
        // points 0..15 are gathered unconditionally (index 0 past the
        // patch's own points) and accumulated only where j < n, so the
        // gathers issue without branches; 16..19 belong to Gregory patches
        int cv[20];
#pragma unroll
        for (int j = 0; j < 20; ++j) cv[j] = j < n ? cvs[j] : 0;
        const float k1 = (float)(1 << depth);

        for (int f0 = 0; f0 < sg.frames; f0 += FC) {
            float acc[3][FC][LC];
#pragma unroll
            for (int o = 0; o < 3; ++o)
#pragma unroll
                for (int q = 0; q < FC; ++q)
#pragma unroll
                    for (int k = 0; k < LC; ++k) acc[o][q][k] = 0.0f;
            auto point = [&](int j) {
                const float *p = sg.src + (size_t)(unsigned)cv[j] * sg.srcV + (size_t)f0 * sg.srcF;
                const bool take = j < n;
#pragma unroll
                for (int q = 0; q < FC; ++q) {
                    if (f0 + q < sg.frames) {                   // uniform per segment
#pragma unroll
                        for (int k = 0; k < LC; ++k) {
                            const float x = p[q * sg.srcF + k];
                            const float a0 = acc[0][q][k] + x * B.p[j];
                            const float a1 = acc[1][q][k] + x * B.ds[j];
                            const float a2 = acc[2][q][k] + x * B.dt[j];
                            acc[0][q][k] = take ? a0 : acc[0][q][k];
                            acc[1][q][k] = take ? a1 : acc[1][q][k];
                            acc[2][q][k] = take ? a2 : acc[2][q][k];
                        }
                    }
                }
            };
#pragma unroll
            for (int j = 0; j < 16; ++j) point(j);
            if (n > 16) {
#pragma unroll
                for (int j = 16; j < 20; ++j) point(j);
            }
#pragma unroll
            for (int q = 0; q < FC; ++q) {
                if (f0 + q >= sg.frames) break;
#pragma unroll
                for (int o = 0; o < 3; ++o) {
                    float *d = sg.out[o];
                    if (!d) continue;
                    const float f = o == 0 ? 1.0f : k1;
                    d += (size_t)i * sg.outV + (size_t)(f0 + q) * sg.outF;
#pragma unroll
                    for (int k = 0; k < LC; ++k) d[k] = acc[o][q][k] * f;
                }
            }
        }
    }
}

}  // namespace osdhip
