// EvalStencils on locality-ordered slices (HipStencilTable layout 2).
//
// The table's rows are re-ordered so that each 64-row slice surrounds a few
// control vertices (rows sorted by their dominant control vertex, possibly
// regrouped by size within small windows of that order), and each
// slice lists its distinct control vertices.  A wavefront evaluating a slice
// first copies those vertices' primvars into LDS -- one vertex per lane,
// element k of slot u at lds[k][u], so a gather of 64 different slots reads
// 64 different banks -- then reads every entry's control point from LDS
// instead of issuing a 32-byte gather to L2 per entry.  Rows keep their
// entries in table order and accumulate w * x from zero, so the sums are
// bit-identical to CpuEvalStencils; each result is written to its Far row.
//
// Padding records of a listed slice point at slot kTileZero, which holds
// zeros: adding w * x = 0 * 0 = +0 leaves every partial sum unchanged (a sum
// that starts at +0 is never -0 under round-to-nearest), so no predicate is
// needed.  A slice with more than kTileZero distinct vertices keeps control
// indices in its records and gathers from memory, with padding records
// (index kTilePad) skipped by a predicate.
#pragma once
#include "common.h"
#include "stencil_sliced.h"

namespace osdhip {

constexpr int kTileSlots = kWave;          // LDS slots per slice
constexpr int kTileZero = kTileSlots - 1;  // the zero slot; at most kTileZero listed vertices
constexpr unsigned kTilePad = 0xffffffffu; // padding record of an unlisted slice

struct TiledArgs {
    const float *src;       // includes srcDesc.offset and the element window start
    int srcStride;
    float *dst[6];          // NULL = output not requested
    int dstStride[6];
    const int *perm;        // slice row -> Far row
    const int *nuniq;
    const long long *ctrlOff;
    const int *ctrl;
    const long long *sliceOff;
    const int *sliceWidth;
    const unsigned *rec;
    int rowBegin, rowEnd;   // Far rows [rowBegin, rowEnd); output row = row - rowBegin
    int nslices;
    int nblocks;
};

template <int NWT, int LC, int G>
__global__ __launch_bounds__(kBlock) void k_stencils_tiled(TiledArgs a)
{
    constexpr int R = RecShape<NWT>::R;
    __shared__ float lds[kWavesPerBlock][LC][kTileSlots];
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int wv = wave_in_block();
    const int lb = xcd_logical_block();
    if (lb >= a.nblocks) return;
    const int s = lb * kWavesPerBlock + wv;
    if (s >= a.nslices) return;                      // whole wave idle
    float(*L)[kTileSlots] = lds[wv];
    const int row = a.perm[(long long)s * kWave + lane];
    const bool active = row >= a.rowBegin && row < a.rowEnd;
    const int W = a.sliceWidth[s];
    const int U = a.nuniq[s];
    const unsigned *rp = a.rec + (size_t)(a.sliceOff[s] + lane) * R;
    const size_t sstride = (size_t)a.srcStride;

    float acc[NWT][LC];
#pragma unroll
    for (int w = 0; w < NWT; ++w)
#pragma unroll
        for (int k = 0; k < LC; ++k) acc[w][k] = 0.0f;

    if (U >= 0) {
        if (lane < U) {
            const float *p = a.src + (size_t)(unsigned)a.ctrl[a.ctrlOff[s] + lane] * sstride;
#pragma unroll
            for (int k = 0; k < LC; ++k) L[k][lane] = p[k];
        } else if (lane == kTileZero) {
#pragma unroll
            for (int k = 0; k < LC; ++k) L[k][lane] = 0.0f;
        }
        wave_lds_sync();
        for (int j = 0; j < W; j += G) {
            int slot[G];
            float ww[NWT][G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                // unconditional loads; past the width: the zero slot
                load_record<NWT, G>(rp + (size_t)min(j + u, W - 1) * (kWave * R), slot[u], ww, u);
                slot[u] = j + u < W ? slot[u] : kTileZero;
#pragma unroll
                for (int w = 0; w < NWT; ++w) ww[w][u] = j + u < W ? ww[w][u] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < G; ++u) {
#pragma unroll
                for (int k = 0; k < LC; ++k) {
                    const float x = L[k][slot[u]];
#pragma unroll
                    for (int w = 0; w < NWT; ++w) acc[w][k] = acc[w][k] + x * ww[w][u];
                }
            }
        }
    } else {
        for (int j = 0; j < W; j += G) {
            int ii[G];
            float ww[NWT][G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                load_record<NWT, G>(rp + (size_t)min(j + u, W - 1) * (kWave * R), ii[u], ww, u);
                ii[u] = j + u < W ? ii[u] : (int)kTilePad;
            }
            float v[G][LC];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const unsigned ix = (unsigned)ii[u] == kTilePad ? 0u : (unsigned)ii[u];
                const float *p = a.src + (size_t)ix * sstride;
#pragma unroll
                for (int k = 0; k < LC; ++k) v[u][k] = p[k];
            }
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const bool take = (unsigned)ii[u] != kTilePad;
#pragma unroll
                for (int w = 0; w < NWT; ++w)
#pragma unroll
                    for (int k = 0; k < LC; ++k) {
                        const float t = acc[w][k] + v[u][k] * ww[w][u];
                        acc[w][k] = take ? t : acc[w][k];
                    }
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

}  // namespace osdhip
