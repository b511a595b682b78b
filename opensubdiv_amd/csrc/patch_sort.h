// Patch-ordered evaluation of PatchCoords: a counting sort of the coords by
// patch (bucket = patchIndex mod 2^20, one patch per bucket up to 2^20
// patches) that produces the order k_patches visits them in.  Each coord's
// result depends on that coord only and goes to its own row, so any visit
// order gives the same bits; in patch order the lanes of a wave share index
// rows and control points in L2 instead of missing on each.
//
// Workspace (device memory, PatchSortLayout): counts[2^20], offsets[2^20],
// 1024 chunk sums, order[n].  Every launch writes what it reads first
// (counts are cleared by a fill kernel, not a memset, so the sequence is
// also safe to capture into a graph), so a workspace may be reused by
// calls that do not overlap.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/osdhip.h"
#include "common.h"
#include "patch_sort_layout.h"

namespace osdhip {

static_assert(kScanChunk == 4 * kBlock, "a scan block takes 4 buckets per thread");

__device__ __forceinline__ unsigned patch_bucket(int patchIndex)
{
    return (unsigned)patchIndex & (unsigned)(kPatchBuckets - 1);
}

__global__ __launch_bounds__(kBlock) void k_zero_u32(unsigned *p, int n)
{
    for (int i = (int)(blockIdx.x * kBlock + threadIdx.x); i < n; i += (int)(gridDim.x * kBlock))
        p[i] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_patch_bucket_count(const osdhip_patch_coord *coords,
                                                               int n, unsigned *counts)
{
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i < n) atomicAdd(&counts[patch_bucket(coords[i].patchIndex)], 1u);
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

// Block b: exclusive scan of counts[b * 1024, +1024) into offsets, and the
// chunk's total into sums[b].
__global__ __launch_bounds__(kBlock) void k_scan_chunks(const unsigned *counts, unsigned *offsets,
                                                        unsigned *sums)
{
    __shared__ unsigned ws[kWavesPerBlock];
    const size_t base = (size_t)blockIdx.x * kScanChunk;
    const unsigned total = block_exclusive_scan4(counts + base, offsets + base, ws);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block: exclusive scan of the 1024 chunk sums in place.
__global__ __launch_bounds__(kBlock) void k_scan_sums(unsigned *sums)
{
    __shared__ unsigned ws[kWavesPerBlock];
    __shared__ unsigned in[kScanChunks];
    for (int i = (int)threadIdx.x; i < kScanChunks; i += kBlock) in[i] = sums[i];
    __syncthreads();
    block_exclusive_scan4(in, sums, ws);
}

__global__ __launch_bounds__(kBlock) void k_scan_add(unsigned *offsets, const unsigned *sums)
{
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    offsets[i] += sums[i / kScanChunk];
}

// Each coord takes the next free position of its bucket (the order within
// a bucket is whatever the atomics give; it does not change any result).
__global__ __launch_bounds__(kBlock) void k_patch_bucket_scatter(const osdhip_patch_coord *coords,
                                                                 int n, unsigned *cursor,
                                                                 int *order)
{
    const int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= n) return;
    const unsigned pos = atomicAdd(&cursor[patch_bucket(coords[i].patchIndex)], 1u);
    order[pos] = i;
}

// Queues the sort of n coords on st in the workspace at ws (laid out by
// PatchSortLayout(n)); returns the order array.
inline int *queue_patch_sort(const osdhip_patch_coord *coords, int n, char *ws, hipStream_t st)
{
    const PatchSortLayout L(n);
    unsigned *counts = (unsigned *)(ws + L.counts), *offsets = (unsigned *)(ws + L.offsets);
    unsigned *sums = (unsigned *)(ws + L.sums);
    int *order = (int *)(ws + L.order);
    const int g = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_zero_u32, dim3(1024), dim3(kBlock), 0, st, counts, kPatchBuckets);
    hipLaunchKernelGGL(k_patch_bucket_count, dim3(g), dim3(kBlock), 0, st, coords, n, counts);
    hipLaunchKernelGGL(k_scan_chunks, dim3(kScanChunks), dim3(kBlock), 0, st, counts, offsets,
                       sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, st, sums);
    hipLaunchKernelGGL(k_scan_add, dim3(kPatchBuckets / kBlock), dim3(kBlock), 0, st, offsets,
                       sums);
    hipLaunchKernelGGL(k_patch_bucket_scatter, dim3(g), dim3(kBlock), 0, st, coords, n, offsets,
                       order);
    return order;
}

}  // namespace osdhip
