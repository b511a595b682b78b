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
    const int sl = lb * kWavesPerBlock + wave_in_block();
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
            // columns past the width re-read the last one (never accumulated:
            // j + u >= W >= size), so the loads are unconditional
            const long long e = eo + (long long)min(j + u, W - 1) * kWave;
            ii[u] = __builtin_nontemporal_load(a.idx + e);
#pragma unroll
            for (int w = 0; w < NW; ++w) ww[w][u] = __builtin_nontemporal_load(a.w[w] + e);
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

// ---------------------------------------------------------------------------
// Record-interleaved slices (the default HipStencilTable layout).
//
// Same 64-row slices, but entry j of row r is ONE record of R dwords,
// {index, w0 [, w1, w2 [, w3, w4, w5, pad]]} (R = 2 / 4 / 8 for 1 / 3 / 6
// weight arrays), stored at (sliceOffset[s] + 64*j + r%64) * R.  Lane r reads
// its whole entry with one 8- or 16-byte load and a wave reads one
// contiguous 64*R*4-byte block per column: a single sequential stream per
// wave instead of 1 + NW streams.  Padding records (j >= size, and rows past
// the table end) hold index 0 and zero weights, so every gather address is
// valid and the gathers are issued without branches; only the accumulation
// is predicated (j < size), which keeps the sums bit-identical to
// CpuEvalStencils.

template <int NWT> struct RecShape;
template <> struct RecShape<1> { static constexpr int R = 2; };
template <> struct RecShape<3> { static constexpr int R = 4; };
template <> struct RecShape<6> { static constexpr int R = 8; };

struct RecArgs {
    const float *src;       // includes srcDesc.offset and the element window start
    int srcStride;
    float *dst[6];          // NULL = output not requested
    int dstStride[6];
    const int *rowSize;     // per slice row (padded to a multiple of 64)
    const long long *sliceOff;
    const int *sliceWidth;
    const unsigned *rec;
    int rowBegin, rowEnd;   // stencils [rowBegin, rowEnd); output row = row - rowBegin
    int sliceBegin, nslices;
    int nblocks;
};

template <int NWT, int G>
__device__ __forceinline__ void load_record(const unsigned *p, int &idx, float (&w)[NWT][G], int u)
{
    if constexpr (NWT == 1) {
        const v2i q = __builtin_nontemporal_load((const v2i *)p);
        idx = q.x;
        w[0][u] = __int_as_float(q.y);
    } else if constexpr (NWT == 3) {
        const v4i q = __builtin_nontemporal_load((const v4i *)p);
        idx = q.x;
        w[0][u] = __int_as_float(q.y);
        w[1][u] = __int_as_float(q.z);
        w[2][u] = __int_as_float(q.w);
    } else {
        const v4i q = __builtin_nontemporal_load((const v4i *)p);
        const v4i r = __builtin_nontemporal_load((const v4i *)p + 1);
        idx = q.x;
        w[0][u] = __int_as_float(q.y);
        w[1][u] = __int_as_float(q.z);
        w[2][u] = __int_as_float(q.w);
        w[3][u] = __int_as_float(r.x);
        w[4][u] = __int_as_float(r.y);
        w[5][u] = __int_as_float(r.z);
    }
}

template <int NWT, int LC, int G>
__global__ __launch_bounds__(kBlock) void k_stencils_rec(RecArgs a)
{
    constexpr int R = RecShape<NWT>::R;
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int lb = xcd_logical_block();
    if (lb >= a.nblocks) return;
    const int sl = lb * kWavesPerBlock + wave_in_block();
    if (sl >= a.nslices) return;                     // whole wave idle
    const int s = a.sliceBegin + sl;
    const int row = s * kWave + lane;
    const bool active = row >= a.rowBegin && row < a.rowEnd;
    const int sz = active ? a.rowSize[row] : 0;
    const int W = a.sliceWidth[s];
    const unsigned *rp = a.rec + (size_t)(a.sliceOff[s] + lane) * R;

    float acc[NWT][LC];
#pragma unroll
    for (int w = 0; w < NWT; ++w)
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[w][k] = 0.0f;

    const float *__restrict__ src = a.src;
    const size_t sstride = (size_t)a.srcStride;
    for (int j = 0; j < W; j += G) {
        int ii[G];
        float ww[NWT][G];
#pragma unroll
        for (int u = 0; u < G; ++u)       // past the width: the last column again, not accumulated
            load_record<NWT, G>(rp + (size_t)min(j + u, W - 1) * (kWave * R), ii[u], ww, u);
        float v[G][LC];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const float *p = src + (size_t)(unsigned)ii[u] * sstride;
#pragma unroll
            for (int k = 0; k < LC; ++k) v[u][k] = p[k];
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const bool take = j + u < sz;
#pragma unroll
            for (int w = 0; w < NWT; ++w)
#pragma unroll
                for (int k = 0; k < LC; ++k) {
                    const float t = acc[w][k] + v[u][k] * ww[w][u];
                    acc[w][k] = take ? t : acc[w][k];
                }
        }
    }
    if (active) {
        const size_t r = (size_t)(row - a.rowBegin);
#pragma unroll
        for (int w = 0; w < NWT; ++w) {
            float *d = a.dst[w];
            if (!d) continue;
            d += r * (size_t)a.dstStride[w];
#pragma unroll
            for (int k = 0; k < LC; ++k) d[k] = acc[w][k];
        }
    }
}

// One weight array, paired records (layout 3): a lane loads the records of
// entries 2p and 2p + 1 of its row with one 16-byte load, so a wave streams
// 1 KiB of records per load instruction (the 8-byte records of layout 0
// give half that).  G entries (G / 2 pair loads) are loaded, then all their
// control-point gathers are issued; past the width the last pair is
// re-read and not accumulated.  Same products and order as CpuEvalStencils.
template <int LC, int G>
__global__ __launch_bounds__(kBlock) void k_stencils_rec_paired(RecArgs a)
{
    static_assert(G % 2 == 0, "pairs");
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int lb = xcd_logical_block();
    if (lb >= a.nblocks) return;
    const int sl = lb * kWavesPerBlock + wave_in_block();
    if (sl >= a.nslices) return;                     // whole wave idle
    const int s = a.sliceBegin + sl;
    const int row = s * kWave + lane;
    const bool active = row >= a.rowBegin && row < a.rowEnd;
    const int sz = active ? a.rowSize[row] : 0;
    const int W = a.sliceWidth[s];
    const int P = (W + 1) >> 1;                      // column pairs
    const unsigned *rp = a.rec + (size_t)(2 * a.sliceOff[s] + 4 * lane);

    float acc[LC];
#pragma unroll
    for (int k = 0; k < LC; ++k) acc[k] = 0.0f;
    const float *__restrict__ src = a.src;
    const size_t sstride = (size_t)a.srcStride;
    for (int p0 = 0; p0 < P; p0 += G / 2) {
        int ii[G];
        float ww[G];
#pragma unroll
        for (int u = 0; u < G / 2; ++u) {
            const v4i q = __builtin_nontemporal_load(
                (const v4i *)(rp + (size_t)min(p0 + u, P - 1) * (kWave * 4)));
            ii[2 * u] = q.x;
            ww[2 * u] = __int_as_float(q.y);
            ii[2 * u + 1] = q.z;
            ww[2 * u + 1] = __int_as_float(q.w);
        }
        float v[G][LC];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const float *pp = src + (size_t)(unsigned)ii[u] * sstride;
#pragma unroll
            for (int k = 0; k < LC; ++k) v[u][k] = pp[k];
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const bool take = 2 * p0 + u < sz;
#pragma unroll
            for (int k = 0; k < LC; ++k) {
                const float t = acc[k] + v[u][k] * ww[u];
                acc[k] = take ? t : acc[k];
            }
        }
    }
    if (active) {
        float *d = a.dst[0];
        if (d) {
            d += (size_t)(row - a.rowBegin) * (size_t)a.dstStride[0];
#pragma unroll
            for (int k = 0; k < LC; ++k) d[k] = acc[k];
        }
    }
}

}  // namespace osdhip
