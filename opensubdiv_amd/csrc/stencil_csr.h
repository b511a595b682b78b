// EvalStencils on a raw CSR stencil table (sizes / offsets / indices /
// weights as Far lays them out, osd/cpuKernel.cpp:71-240).
//
// One 64-lane wavefront owns 64 consecutive stencils.  Their table entries
// form one slice [lo, hi) of the index / weight streams, which the wave
// copies into LDS with coalesced, non-temporal 16-byte loads (all loads of
// the slice are issued before the first is consumed, so each wave keeps up
// to ~10 KB of HBM traffic in flight).  Each lane then reduces its own
// stencil from LDS, issuing four control-point gathers at a time through L2.
// Entries are accumulated in table order without FMA contraction, which
// reproduces CpuEvalStencils bit for bit.
#pragma once
#include <climits>

#include "common.h"

namespace osdhip {

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

constexpr int kGatherGroup = 4;

// Accumulates entries [j0, sz) of one stencil; idxAt(j) / wAt(w, j) fetch
// the j-th entry.  Gathers of a group are issued before any is consumed.
template <int NW, int LC, class IDX, class WT>
__device__ __forceinline__ void accumulate_stencil(const float *__restrict__ src, size_t sstride,
                                                   int sz, IDX idxAt, WT wAt,
                                                   float (&acc)[NW][LC])
{
    for (int j = 0; j < sz; j += kGatherGroup) {
        float v[kGatherGroup][LC];
#pragma unroll
        for (int u = 0; u < kGatherGroup; ++u) {
            if (j + u < sz) {
                const float *p = src + (size_t)(unsigned)idxAt(j + u) * sstride;
#pragma unroll
                for (int k = 0; k < LC; ++k) v[u][k] = p[k];
            }
        }
#pragma unroll
        for (int u = 0; u < kGatherGroup; ++u) {
            if (j + u < sz) {
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const float ww = wAt(w, j + u);
#pragma unroll
                    for (int k = 0; k < LC; ++k) acc[w][k] += v[u][k] * ww;
                }
            }
        }
    }
}

// NW weight arrays (1: P, 3: P/du/dv, 6: + duu/duv/dvv); LC floats per vertex
// in this launch; CAP = table entries staged per wave; VEC = index and weight
// pointers are 16-byte aligned (16-byte staging loads).
template <int NW, int LC, int CAP, bool VEC>
__global__ __launch_bounds__(kBlock) void k_stencils_csr(StencilArgs a)
{
    constexpr int SLOTS = CAP + 4;          // room for the 16-byte alignment slack
    __shared__ __attribute__((aligned(16))) int s_idx[kWavesPerBlock][SLOTS];
    __shared__ __attribute__((aligned(16))) float s_w[kWavesPerBlock][NW][SLOTS];

    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wv = wave_in_block();
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
    const int base = VEC ? (lo & ~3) : lo;
    const int span = hi - base;

    float acc[NW][LC];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[w][k] = 0.0f;

    const float *__restrict__ src = a.src;
    const size_t sstride = (size_t)a.srcStride;

    if (span <= SLOTS) {
        if constexpr (VEC) {
            // [base, base + 4*lane + 256*it) chunks; a chunk is loaded as one
            // 16-byte access when it lies inside [base, hi), else per entry.
            constexpr int IT = (SLOTS + 4 * kWave - 1) / (4 * kWave);
            v4i ri[IT];
            v4f rw[NW][IT];
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int e = it * 4 * kWave + 4 * lane;
                if (e + 4 <= span) {
                    ri[it] = __builtin_nontemporal_load((const v4i *)(a.indices + base + e));
#pragma unroll
                    for (int w = 0; w < NW; ++w)
                        rw[w][it] = __builtin_nontemporal_load((const v4f *)(a.w[w] + base + e));
                } else if (e < span) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const bool in = e + q < span;
                        ri[it][q] = in ? a.indices[base + e + q] : 0;
#pragma unroll
                        for (int w = 0; w < NW; ++w) rw[w][it][q] = in ? a.w[w][base + e + q] : 0.0f;
                    }
                }
            }
#pragma unroll
            for (int it = 0; it < IT; ++it) {
                const int e = it * 4 * kWave + 4 * lane;
                if (e < span) {
                    *(v4i *)&s_idx[wv][e] = ri[it];
#pragma unroll
                    for (int w = 0; w < NW; ++w) *(v4f *)&s_w[wv][w][e] = rw[w][it];
                }
            }
        } else {
            for (int e = lane; e < span; e += kWave) {
                s_idx[wv][e] = __builtin_nontemporal_load(a.indices + base + e);
#pragma unroll
                for (int w = 0; w < NW; ++w)
                    s_w[wv][w][e] = __builtin_nontemporal_load(a.w[w] + base + e);
            }
        }
        wave_lds_sync();
        const int rb = off - base;
        accumulate_stencil<NW, LC>(
            src, sstride, sz, [&](int j) { return s_idx[wv][rb + j]; },
            [&](int w, int j) { return s_w[wv][w][rb + j]; }, acc);
    } else {
        // Slice larger than the staging buffer: read the table directly.
        accumulate_stencil<NW, LC>(
            src, sstride, sz, [&](int j) { return a.indices[off + j]; },
            [&](int w, int j) { return a.w[w][off + j]; }, acc);
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

}  // namespace osdhip
