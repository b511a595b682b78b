// Workspace layout of the patch-order binning (patch_sort.h): the per-block
// bin histograms hist[kPatchBins x blocks] (bin-major; scanned in place into
// bin offsets), the scan's chunk sums, and order[n]; 256-byte aligned.
#pragma once
#include <cstddef>

namespace osdhip {

constexpr int kPatchBins = 1024;          // bin = (patchIndex >> 10) mod 1024
constexpr int kPatchBinShift = 10;        // 1024 consecutive patches per bin
constexpr int kScanChunk = 1024;          // histogram entries per scan block
constexpr int kMaxBinBlocks = 1024;       // blocks (tiles of coords) at most

// Coords per block: at least 4096, and few enough blocks that the
// histograms fit one two-level scan (kMaxBinBlocks chunks of kScanChunk).
constexpr long long patch_tile(long long n)
{
    long long t = (n + kMaxBinBlocks - 1) / kMaxBinBlocks;
    t = (t + 255) / 256 * 256;
    return t < 4096 ? 4096 : t;
}

struct PatchSortLayout {
    size_t hist = 0, sums = 0, order = 0, bytes = 0;
    int tile = 0, blocks = 0;
    constexpr explicit PatchSortLayout(long long n)
    {
        auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
        tile = (int)patch_tile(n);
        blocks = (int)((n + tile - 1) / tile);
        sums = up(hist + 4 * (size_t)kPatchBins * (size_t)(blocks > 0 ? blocks : 1));
        order = up(sums + 4 * (size_t)kMaxBinBlocks);
        bytes = up(order + 4 * (size_t)n);
    }
};

}  // namespace osdhip
