// Workspace layout of the patch-order counting sort (patch_sort.h):
// counts[2^20], offsets[2^20], 1024 chunk sums, order[n], 256-byte aligned.
#pragma once
#include <cstddef>

namespace osdhip {

constexpr int kPatchBuckets = 1 << 20;
constexpr int kScanChunk = 1024;                    // buckets per scan block
constexpr int kScanChunks = kPatchBuckets / kScanChunk;   // 1024
static_assert(kScanChunks == kScanChunk, "one block scans the chunk sums");

struct PatchSortLayout {
    size_t counts = 0, offsets = 0, sums = 0, order = 0, bytes = 0;
    constexpr explicit PatchSortLayout(long long n)
    {
        auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
        offsets = up(counts + 4 * (size_t)kPatchBuckets);
        sums = up(offsets + 4 * (size_t)kPatchBuckets);
        order = up(sums + 4 * (size_t)kScanChunks);
        bytes = up(order + 4 * (size_t)n);
    }
};

}  // namespace osdhip
