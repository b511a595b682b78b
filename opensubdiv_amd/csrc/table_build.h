// Device-side construction of HipStencilTable's slice layouts from the
// uploaded CSR arrays (used by osdhip_stencil_table_create only).
#pragma once
#include "common.h"
#include "stencil_sliced.h"
#include "stencil_tiled.h"

namespace osdhip {

// Per slice: widest stencil, and 64 * width entries (the scan input).
__global__ __launch_bounds__(kBlock) void k_slice_widths(const int *sizes, int n, int S,
                                                         int *width, long long *entries,
                                                         int paired = 0)
{
    const int s = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (s >= S) return;
    int w = 0;
    const int r1 = min(n, (s + 1) * kWave);
    for (int r = s * kWave; r < r1; ++r) w = max(w, sizes[r]);
    width[s] = w;
    // paired records: a slice holds whole column pairs
    entries[s] = (long long)(paired ? w + (w & 1) : w) * kWave;
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

// Paired records (one weight array): entries 2p and 2p + 1 of a row sit
// side by side, {index, w, index, w}, so a lane reads both with one 16-byte
// load; column pair p of a slice holds its 64 rows' pairs contiguously,
// record (j, lane) at dword 2 * sliceOff + 256 * (j / 2) + 4 * lane +
// 2 * (j % 2).  An odd width is padded to a whole pair.
__global__ __launch_bounds__(kBlock) void k_fill_records_paired(
    const int *sizes, const int *offsets, const int *indices, const float *w0, int n, int rows,
    const long long *sliceOff, const int *sliceWidth, unsigned *rec)
{
    const int r = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (r >= rows) return;
    const int s = r / kWave, lane = r % kWave;
    const int W = sliceWidth[s], W2 = W + (W & 1);
    const int sz = r < n ? sizes[r] : 0;
    const long long off = r < n ? (long long)offsets[r] : 0;
    for (int j = 0; j < W2; ++j) {
        unsigned q0 = 0u, q1 = 0u;
        if (j < sz) {
            q0 = (unsigned)indices[off + j];
            q1 = __float_as_uint(w0[off + j]);
        }
        unsigned *d = rec + (size_t)(2 * sliceOff[s] + 256LL * (j >> 1) + 4 * lane + 2 * (j & 1));
        d[0] = q0;
        d[1] = q1;
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

// ---------------------------------------------------------------------------
// Locality-ordered slices (layout 2).
//
// Rows are ordered by their dominant control vertex (the entry of largest
// |weight|, the first one on ties; a stable sort keeps Far order among
// equal keys), so the 64 rows of a slice surround a few control vertices and
// share most of their support: 26-38 distinct control vertices per slice on
// creased cube-spheres at levels 2-4, against 30 (level 4) to 100 (level 2)
// on average in Far order -- which keeps a slice's 64 outputs consecutive
// (order option 1).  Per slice
// the distinct control vertices are listed in ascending order; when there
// are at most kTileZero of them the slice's records hold list positions
// (LDS slots), otherwise control indices.

// Sort key per slice row: the dominant control vertex (empty rows:
// 0xffffffff), or with farOrder the row itself (Far's order kept); rows
// past n: 0xffffffff, after every real key; value: the row.
__global__ __launch_bounds__(kBlock) void k_row_dominant(const int *sizes, const int *offsets,
                                                         const int *indices, const float *w0, int n,
                                                         int rows, int farOrder, unsigned *keys,
                                                         int *vals)
{
    const int r = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (r >= rows) return;
    unsigned key = 0xffffffffu;
    if (farOrder) {
        if (r < n) key = (unsigned)r;
    } else if (r < n && sizes[r] > 0) {
        const long long off = offsets[r];
        key = (unsigned)indices[off];
        float best = fabsf(w0[off]);
        for (int j = 1; j < sizes[r]; ++j) {
            const float a = fabsf(w0[off + j]);
            if (a > best) {
                best = a;
                key = (unsigned)indices[off + j];
            }
        }
    }
    keys[r] = key;
    vals[r] = r;
}

// Rows of the dominant-vertex order regrouped by size within windows of K
// consecutive rows (one block per window): a slice costs its widest row in
// every lane, and rows of one neighbourhood differ in size (9, 12, 16
// entries at a regular vertex, edge and face), so sorting each window's rows
// by size -- stable, by counting -- makes a slice's rows alike while its
// rows still come from a few neighbouring control vertices.  Rows past n
// (slice padding) stay after every real row.
template <int K>
__global__ __launch_bounds__(kBlock) void k_window_size_sort(const int *in, const int *sizes,
                                                             int n, long long rows, int *out)
{
    __shared__ unsigned char key[K];
    __shared__ int start[65];
    const long long w0 = (long long)blockIdx.x * K;
    const int m = (int)(rows - w0 < K ? rows - w0 : K);
    for (int b = (int)threadIdx.x; b < 65; b += kBlock) start[b] = 0;
    for (int i = (int)threadIdx.x; i < m; i += kBlock) {
        const int r = in[w0 + i];
        key[i] = (unsigned char)(r >= 0 && r < n ? min(sizes[r], 63) : 64);
    }
    __syncthreads();
    if (threadIdx.x == 0) {                    // counts, then their exclusive sums
        for (int i = 0; i < m; ++i) ++start[key[i]];
        int s = 0;
        for (int b = 0; b < 65; ++b) {
            const int c = start[b];
            start[b] = s;
            s += c;
        }
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < m; i += kBlock) {
        const int k = key[i];
        int before = 0;
        for (int j = 0; j < i; ++j) before += key[j] == k;
        out[w0 + start[k] + before] = in[w0 + i];
    }
}

// One wave per slice of the ordered rows perm[].  Count pass (FILL false):
// slice width, record entries (64 * width), distinct control vertices
// (nuniq = -1 when more than kTileZero) and list length.  Fill pass: the
// ascending control list and the records -- {list position, weights} for a
// listed slice (padding: {kTileZero, 0}), {control index, weights} otherwise
// (padding: {kTilePad, 0}).  Distinct vertices are extracted in ascending
// order by repeated wave minima over the slice's entries.
template <int NWT, bool FILL>
__global__ __launch_bounds__(kBlock) void k_tiled_slices(
    const int *perm, const int *sizes, const int *offsets, const int *indices, const float *w0,
    const float *w1, const float *w2, const float *w3, const float *w4, const float *w5, int n, int S,
    int *width, long long *entries, int *nuniq, long long *listLen, const long long *sliceOff,
    const long long *ctrlOff, int *ctrl, unsigned *rec)
{
    constexpr int R = RecShape<NWT>::R;
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int s = (int)blockIdx.x * kWavesPerBlock + wave_in_block();
    if (s >= S) return;                                // whole wave
    const int r = perm[(long long)s * kWave + lane];
    const bool valid = r >= 0 && r < n;
    const int sz = valid ? sizes[r] : 0;
    const long long off = valid ? (long long)offsets[r] : 0;
    const float *ws[6] = { w0, w1, w2, w3, w4, w5 };
    const int W = FILL ? width[s] : wave_max_i(sz);
    const bool listed = FILL ? nuniq[s] >= 0 : true;
    unsigned *rp = FILL ? rec + (size_t)(sliceOff[s] + lane) * R : nullptr;
    int U = 0;
    bool over = false;
    if (listed) {
        long long last = -1;
        for (;;) {
            unsigned m = 0xffffffffu;
            for (int j = 0; j < sz; ++j) {
                const unsigned x = (unsigned)indices[off + j];
                if ((long long)x > last) m = min(m, x);
            }
            m = wave_min_u(m);
            if (m == 0xffffffffu) break;
            if (U == kTileZero) {
                over = true;
                break;
            }
            if (FILL) {
                if (lane == 0) ctrl[ctrlOff[s] + U] = (int)m;
                for (int j = 0; j < sz; ++j) {
                    if ((unsigned)indices[off + j] != m) continue;
                    unsigned *d = rp + (size_t)j * (kWave * R);
                    d[0] = (unsigned)U;
#pragma unroll
                    for (int k = 0; k < NWT; ++k) d[1 + k] = __float_as_uint(ws[k][off + j]);
#pragma unroll
                    for (int k = 1 + NWT; k < R; ++k) d[k] = 0u;
                }
            }
            last = m;
            ++U;
        }
    }
    if (!FILL) {
        if (lane == 0) {
            width[s] = W;
            entries[s] = (long long)W * kWave;
            nuniq[s] = over ? -1 : U;
            listLen[s] = over ? 0 : U;
        }
        return;
    }
    for (int j = 0; j < W; ++j) {
        unsigned *d = rp + (size_t)j * (kWave * R);
        if (listed && j < sz) continue;                // written above
        unsigned q[R];
#pragma unroll
        for (int k = 0; k < R; ++k) q[k] = 0u;
        if (j < sz) {
            q[0] = (unsigned)indices[off + j];
#pragma unroll
            for (int k = 0; k < NWT; ++k) q[1 + k] = __float_as_uint(ws[k][off + j]);
        } else {
            q[0] = listed ? (unsigned)kTileZero : kTilePad;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) d[k] = q[k];
    }
}

}  // namespace osdhip
