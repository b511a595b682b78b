// EvalStencils on the device layout of HipStencilTable: 64-row slices.
//
// HipStencilTable::Create re-lays the Far CSR table out for wave64: the
// entries of the 64 stencils of slice s are stored column-major, entry j of
// row r at sliceOffset[s] + 64*j + (r % 64), padded to the slice's widest
// stencil.  Lane r of a wavefront then reads entry j of its own stencil
// with one fully coalesced 256-byte load per stream, without LDS (so the
// kernel runs at full occupancy), and accumulates entries in table order
// (bit-identical to CpuEvalStencils).  Padding entries are loaded but never
// accumulated (per-lane predicate j < size).  The control-point gathers of
// kSliceGroup consecutive entries are issued before any is consumed.
#pragma once
#include "common.h"

namespace osdhip {

struct SlicedArgs {
    const float *src;       // includes srcDesc.offset and the element window start
    int srcStride;
    float *dst[6];
    int dstStride[6];
    const int *rowSize;     // per stencil
    const long long *sliceOff;
    const int *sliceWidth;
    const int *idx;
    const float *w[6];
    int rowBegin, rowEnd;   // stencils [rowBegin, rowEnd); output row = row - rowBegin
    int sliceBegin, nslices;
    int nblocks;
};

constexpr int kSliceGroup = 4;

template <int NW, int LC>
__global__ __launch_bounds__(kBlock) void k_stencils_sliced(SlicedArgs a)
{
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int lb = xcd_logical_block();
    if (lb >= a.nblocks) return;
    const int sl = lb * kWavesPerBlock + (int)threadIdx.x / kWave;
    if (sl >= a.nslices) return;                     // whole wave idle
    const int s = a.sliceBegin + sl;
    const int row = s * kWave + lane;
    const bool active = row >= a.rowBegin && row < a.rowEnd;
    const int sz = active ? a.rowSize[row] : 0;
    const int W = a.sliceWidth[s];
    const long long eo = a.sliceOff[s] + lane;

    float acc[NW][LC];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[w][k] = 0.0f;

    const float *__restrict__ src = a.src;
    const size_t sstride = (size_t)a.srcStride;
    for (int j = 0; j < W; j += kSliceGroup) {
        int ii[kSliceGroup];
        float ww[NW][kSliceGroup];
#pragma unroll
        for (int u = 0; u < kSliceGroup; ++u) {
            if (j + u < W) {                          // wave-uniform
                const long long e = eo + (long long)(j + u) * kWave;
                ii[u] = __builtin_nontemporal_load(a.idx + e);
#pragma unroll
                for (int w = 0; w < NW; ++w) ww[w][u] = __builtin_nontemporal_load(a.w[w] + e);
            }
        }
        float v[kSliceGroup][LC];
#pragma unroll
        for (int u = 0; u < kSliceGroup; ++u) {
            if (j + u < sz) {
                const float *p = src + (size_t)(unsigned)ii[u] * sstride;
#pragma unroll
                for (int k = 0; k < LC; ++k) v[u][k] = p[k];
            }
        }
#pragma unroll
        for (int u = 0; u < kSliceGroup; ++u) {
            if (j + u < sz) {
#pragma unroll
                for (int w = 0; w < NW; ++w)
#pragma unroll
                    for (int k = 0; k < LC; ++k) acc[w][k] += v[u][k] * ww[w][u];
            }
        }
    }
    if (active) {
        const size_t r = (size_t)(row - a.rowBegin);
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            float *d = a.dst[w];
            if (!d) continue;
            d += r * (size_t)a.dstStride[w];
#pragma unroll
            for (int k = 0; k < LC; ++k) d[k] = acc[w][k];
        }
    }
}

}  // namespace osdhip
