// EvalStencils on a raw CSR stencil table (sizes / offsets / indices /
// weights as Far lays them out, osd/cpuKernel.cpp:71-240).
//
// One 64-lane wavefront owns 64 consecutive stencils.  Their table entries
// form one slice [lo, hi) of the index / weight streams, which the wave
// walks in windows of kCsrWindow (512) entries: each window is copied into LDS
// with coalesced, non-temporal 16-byte loads, then every lane adds the
// entries of its own stencil that fall in the window, issuing eight
// control-point gathers at a time through L2.  Small windows keep the LDS
// and register footprint per wave low, so many waves per CU hide the
// loads' latency.  Slice entry e of a window sits at LDS position
// e + e / 16: lane l reading entry j of a stencil that starts 16 l entries
// in (stencils of 16 entries) then hits bank 17 l + j, a different one for
// every lane.
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

constexpr int kGatherGroup = 8;

// LDS position of slice entry e (one padding slot after every 16 entries).
__device__ __forceinline__ int csr_lds_pos(int e) { return e + (e >> 4); }

// Accumulates entries [j0, sz) of one stencil; idxAt(j) / wAt(w, j) fetch
// the j-th entry.  Gathers of a group are issued before any is consumed.
template <int NW, int LC, class IDX, class WT>
__device__ __forceinline__ void accumulate_stencil(const float *__restrict__ src, size_t sstride,
                                                   int j0, int sz, IDX idxAt, WT wAt,
                                                   float (&acc)[NW][LC])
{
    for (int j = j0; j < sz; j += kGatherGroup) {
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

// 16-byte chunks per lane in one LDS window of a wave's table slice, and
// the window's entries.
constexpr int kCsrChunks = 2;
constexpr int kCsrWindow = 4 * kWave * kCsrChunks;

// NW weight arrays (1: P, 3: P/du/dv, 6: + duu/duv/dvv); LC floats per vertex
// in this launch; VEC = index and weight pointers are 16-byte aligned
// (16-byte staging loads).
template <int NW, int LC, bool VEC>
__global__ __launch_bounds__(kBlock) void k_stencils_csr(StencilArgs a)
{
    constexpr int PADDED = kCsrWindow + kCsrWindow / 16;
    __shared__ int s_idx[kWavesPerBlock][PADDED];
    __shared__ float s_w[kWavesPerBlock][NW][PADDED];

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
    lo = __builtin_amdgcn_readfirstlane(lo);
    hi = __builtin_amdgcn_readfirstlane(hi);

    float acc[NW][LC];
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[w][k] = 0.0f;

    const float *__restrict__ src = a.src;
    const size_t sstride = (size_t)a.srcStride;
    const long long base = VEC ? ((long long)lo & ~3LL) : lo;
    for (long long w0 = base; w0 < hi; w0 += kCsrWindow) {
        // window [w0, w0 + kCsrWindow) of the slice; lane l stages entries
        // w0 + 4 (l + 64 c) .. + 3 for chunks c
        v4i ri[kCsrChunks];
        v4f rw[NW][kCsrChunks];
#pragma unroll
        for (int c = 0; c < kCsrChunks; ++c) {
            const long long e = w0 + 4 * (lane + kWave * c);
            if (VEC && e + 4 <= hi) {
                ri[c] = __builtin_nontemporal_load((const v4i *)(a.indices + e));
#pragma unroll
                for (int w = 0; w < NW; ++w)
                    rw[w][c] = __builtin_nontemporal_load((const v4f *)(a.w[w] + e));
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool in = e + q >= lo && e + q < hi;
                    ri[c][q] = in ? a.indices[e + q] : 0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) rw[w][c][q] = in ? a.w[w][e + q] : 0.0f;
                }
            }
        }
        wave_lds_sync();                             // the previous window's reads are done
        // dword stores (a chunk's positions need not be 16-byte aligned)
#pragma unroll
        for (int c = 0; c < kCsrChunks; ++c) {
            const int pe = csr_lds_pos(4 * (lane + kWave * c));
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                s_idx[wv][pe + q] = ri[c][q];
#pragma unroll
                for (int w = 0; w < NW; ++w) s_w[wv][w][pe + q] = rw[w][c][q];
            }
        }
        wave_lds_sync();
        // this lane's entries j in [w0, w0 + kCsrWindow), in table order
        const long long rb = (long long)off - w0;   // window position of entry 0
        const int j0 = (int)max(0LL, -rb);
        const int j1 = (int)min((long long)sz, (long long)kCsrWindow - rb);
        if (j0 < j1)
            accumulate_stencil<NW, LC>(
                src, sstride, j0, j1, [&](int j) { return s_idx[wv][csr_lds_pos((int)rb + j)]; },
                [&](int w, int j) { return s_w[wv][w][csr_lds_pos((int)rb + j)]; }, acc);
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
