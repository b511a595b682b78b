// Device-side construction of HipStencilTable's slice layouts from the
// uploaded CSR arrays (used by osdhip_stencil_table_create only).
#pragma once
#include "common.h"
#include "stencil_sliced.h"

namespace osdhip {

// Per slice: widest stencil, and 64 * width entries (the scan input).
__global__ __launch_bounds__(kBlock) void k_slice_widths(const int *sizes, int n, int S,
                                                         int *width, long long *entries)
{
    const int s = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (s >= S) return;
    int w = 0;
    const int r1 = min(n, (s + 1) * kWave);
    for (int r = s * kWave; r < r1; ++r) w = max(w, sizes[r]);
    width[s] = w;
    entries[s] = (long long)w * kWave;
}

// Slice-row sizes padded to whole slices (rows past the end: 0).
__global__ __launch_bounds__(kBlock) void k_pad_sizes(const int *sizes, int n, int rows,
                                                      int *padded)
{
    const int r = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (r >= rows) return;
    padded[r] = r < n ? sizes[r] : 0;
}

// One lane per slice row: writes the row's W records (padding included).
template <int NWT>
__global__ __launch_bounds__(kBlock) void k_fill_records(
    const int *sizes, const int *offsets, const int *indices, const float *w0, const float *w1,
    const float *w2, const float *w3, const float *w4, const float *w5, int n, int rows,
    const long long *sliceOff, const int *sliceWidth, unsigned *rec)
{
    constexpr int R = RecShape<NWT>::R;
    const int r = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (r >= rows) return;
    const int s = r / kWave, lane = r % kWave;
    const int W = sliceWidth[s];
    const int sz = r < n ? sizes[r] : 0;
    const long long off = r < n ? (long long)offsets[r] : 0;
    const float *ws[6] = { w0, w1, w2, w3, w4, w5 };
    for (int j = 0; j < W; ++j) {
        unsigned q[R];
#pragma unroll
        for (int k = 0; k < R; ++k) q[k] = 0u;
        if (j < sz) {
            q[0] = (unsigned)indices[off + j];
#pragma unroll
            for (int k = 0; k < NWT; ++k) q[1 + k] = __float_as_uint(ws[k][off + j]);
        }
        unsigned *d = rec + (size_t)(sliceOff[s] + (long long)j * kWave + lane) * R;
#pragma unroll
        for (int k = 0; k < R; ++k) d[k] = q[k];
    }
}

// Split-stream layout (index and each weight array in its own column-major
// slice stream), kept as an alternative for the autotuner.
__global__ __launch_bounds__(kBlock) void k_fill_split(
    const int *sizes, const int *offsets, const int *indices, const float *w0, const float *w1,
    const float *w2, const float *w3, const float *w4, const float *w5, int nw, int n, int rows,
    const long long *sliceOff, const int *sliceWidth, int *sidx, float *s0, float *s1, float *s2,
    float *s3, float *s4, float *s5)
{
    const int r = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (r >= rows) return;
    const int s = r / kWave, lane = r % kWave;
    const int W = sliceWidth[s];
    const int sz = r < n ? sizes[r] : 0;
    const long long off = r < n ? (long long)offsets[r] : 0;
    const float *ws[6] = { w0, w1, w2, w3, w4, w5 };
    float *ds[6] = { s0, s1, s2, s3, s4, s5 };
    for (int j = 0; j < W; ++j) {
        const size_t e = (size_t)(sliceOff[s] + (long long)j * kWave + lane);
        sidx[e] = j < sz ? indices[off + j] : 0;
        for (int k = 0; k < nw; ++k) ds[k][e] = j < sz ? ws[k][off + j] : 0.0f;
    }
}

}  // namespace osdhip
