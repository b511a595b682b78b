// Patch-ordered evaluation of PatchCoords: the coords are binned by patch
// (bin = (patchIndex >> 10) mod 1024: 1024 consecutive patches per bin,
// which share control points and whose index rows and points fit in one
// XCD's L2) into the order k_patches visits them in.  Each coord's result
// depends on that coord only and goes to its own row, so any visit order
// gives the same bits.
//
// One pass over the coords counts bins per tile of coords in LDS (no
// global atomics), a two-level scan turns the bin-major histograms into
// each (bin, tile)'s first position, and a second pass places each coord
// (LDS atomics per tile).  Workspace (device memory, PatchSortLayout):
// histograms, chunk sums, order[n]; every launch writes what it reads, with
// kernels only (no memset), so the sequence may be captured into a graph
// and a workspace reused by calls that do not overlap.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/osdhip.h"
#include "common.h"
#include "patch_sort_layout.h"

namespace osdhip {

static_assert(kScanChunk == 4 * kBlock, "a scan block takes 4 entries per thread");
static_assert(kMaxBinBlocks <= kScanChunk, "one block scans the chunk sums");
static_assert(kPatchBins % kBlock == 0, "bins are cleared and written kBlock at a time");

__device__ __forceinline__ int patch_bin(int patchIndex)
{
    return patchIndex < 0 ? 0 : (patchIndex >> kPatchBinShift) & (kPatchBins - 1);
}

// Block b: bin counts of coords [b * tile, min(n, (b + 1) * tile)) into
// hist[bin * blocks + b].
__global__ __launch_bounds__(kBlock) void k_patch_bin_hist(const osdhip_patch_coord *coords, int n,
                                                           int tile, int blocks, unsigned *hist)
{
    __shared__ unsigned h[kPatchBins];
    for (int k = (int)threadIdx.x; k < kPatchBins; k += kBlock) h[k] = 0u;
    __syncthreads();
    const long long i0 = (long long)blockIdx.x * tile;
    const long long i1 = i0 + tile < n ? i0 + tile : n;
    for (long long i = i0 + threadIdx.x; i < i1; i += kBlock)
        atomicAdd(&h[patch_bin(coords[i].patchIndex)], 1u);
    __syncthreads();
    for (int k = (int)threadIdx.x; k < kPatchBins; k += kBlock)
        hist[(size_t)k * blocks + blockIdx.x] = h[k];
}

// Exclusive scan of the kBlock * 4 values at v (4 per thread, thread t owns
// v[4t .. 4t+3]) into out; returns the block total to every thread.
__device__ __forceinline__ unsigned block_exclusive_scan4(const unsigned *v, unsigned *out,
                                                          unsigned *waveSums)
{
    const int t = (int)threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    unsigned x0 = v[4 * t], x1 = v[4 * t + 1], x2 = v[4 * t + 2], x3 = v[4 * t + 3];
    const unsigned mine = x0 + x1 + x2 + x3;
    unsigned incl = mine;                               // inclusive scan across the wave
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const unsigned y = (unsigned)__shfl_up((int)incl, d, kWave);
        if (lane >= d) incl += y;
    }
    if (lane == kWave - 1) waveSums[w] = incl;
    __syncthreads();
    unsigned base = 0, total = 0;
#pragma unroll
    for (int k = 0; k < kWavesPerBlock; ++k) {
        const unsigned s = waveSums[k];
        base += k < w ? s : 0u;
        total += s;
    }
    unsigned e = base + incl - mine;
    out[4 * t] = e;
    e += x0;
    out[4 * t + 1] = e;
    e += x1;
    out[4 * t + 2] = e;
    e += x2;
    out[4 * t + 3] = e;
    return total;
}

// Block b: exclusive scan, in place, of v[b * 1024, +1024), and the
// chunk's total into sums[b].
__global__ __launch_bounds__(kBlock) void k_scan_chunks(unsigned *v, unsigned *sums)
{
    __shared__ unsigned ws[kWavesPerBlock];
    __shared__ unsigned in[kScanChunk];
    const size_t base = (size_t)blockIdx.x * kScanChunk;
    for (int k = (int)threadIdx.x; k < kScanChunk; k += kBlock) in[k] = v[base + k];
    __syncthreads();
    const unsigned total = block_exclusive_scan4(in, v + base, ws);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block: exclusive scan of the first `chunks` chunk sums in place.
__global__ __launch_bounds__(kBlock) void k_scan_sums(unsigned *sums, int chunks)
{
    __shared__ unsigned ws[kWavesPerBlock];
    __shared__ unsigned in[kScanChunk];
    for (int k = (int)threadIdx.x; k < kScanChunk; k += kBlock) in[k] = k < chunks ? sums[k] : 0u;
    __syncthreads();
    block_exclusive_scan4(in, in, ws);          // each thread rewrites only what it read
    __syncthreads();
    for (int k = (int)threadIdx.x; k < chunks; k += kBlock) sums[k] = in[k];
}

// Adds each chunk's offset to its entries.
__global__ __launch_bounds__(kBlock) void k_scan_add(unsigned *v, const unsigned *sums)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    v[i] += sums[i / kScanChunk];
}

// Block b places coords [b * tile, ...) of each bin from the bin's first
// free position for this tile (the order within a tile's share of a bin is
// whatever the LDS atomics give; it does not change any result).
__global__ __launch_bounds__(kBlock) void k_patch_bin_scatter(const osdhip_patch_coord *coords,
                                                              int n, int tile, int blocks,
                                                              const unsigned *offsets, int *order)
{
    __shared__ unsigned cur[kPatchBins];
    for (int k = (int)threadIdx.x; k < kPatchBins; k += kBlock)
        cur[k] = offsets[(size_t)k * blocks + blockIdx.x];
    __syncthreads();
    const long long i0 = (long long)blockIdx.x * tile;
    const long long i1 = i0 + tile < n ? i0 + tile : n;
    for (long long i = i0 + threadIdx.x; i < i1; i += kBlock) {
        const unsigned pos = atomicAdd(&cur[patch_bin(coords[i].patchIndex)], 1u);
        order[pos] = (int)i;
    }
}

// Queues the binning of n coords on st in the workspace at ws (laid out by
// PatchSortLayout(n)); returns the order array.
inline int *queue_patch_sort(const osdhip_patch_coord *coords, int n, char *ws, hipStream_t st)
{
    const PatchSortLayout L(n);
    unsigned *hist = (unsigned *)(ws + L.hist), *sums = (unsigned *)(ws + L.sums);
    int *order = (int *)(ws + L.order);
    const int chunks = L.blocks;             // kPatchBins * blocks entries = blocks chunks
    hipLaunchKernelGGL(k_patch_bin_hist, dim3(L.blocks), dim3(kBlock), 0, st, coords, n, L.tile,
                       L.blocks, hist);
    hipLaunchKernelGGL(k_scan_chunks, dim3(chunks), dim3(kBlock), 0, st, hist, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, st, sums, chunks);
    hipLaunchKernelGGL(k_scan_add, dim3(chunks * (kScanChunk / kBlock)), dim3(kBlock), 0, st,
                       hist, sums);
    hipLaunchKernelGGL(k_patch_bin_scatter, dim3(L.blocks), dim3(kBlock), 0, st, coords, n,
                       L.tile, L.blocks, hist, order);
    return order;
}

}  // namespace osdhip
