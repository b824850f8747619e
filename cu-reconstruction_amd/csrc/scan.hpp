// scan.hpp -- device-wide exclusive scan of int32 counts into int64 offsets
// (three launches: block sums, one-workgroup scan of the block sums, apply),
// shared by marching cubes and the differentiable path.
#pragma once

#include "recon_common.hpp"

namespace recon {

// ---- exclusive scan of int32 counts into int64 offsets ---------------------
constexpr int kScanItems = 4;
constexpr int kScanBlock = 256 * kScanItems;

__device__ __forceinline__ long long block_exclusive_scan(long long v, long long* sh,
                                                          long long* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const long long o = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += o;
        __syncthreads();
    }
    const long long incl = sh[t];
    *total = sh[255];
    __syncthreads();
    return incl - v;
}

// Exclusive scan of n int32 counts into int64 offsets on stream s: three
// launches (block sums, top-level scan, apply); sums[nb] receives the total,
// nb = (n + kScanBlock - 1) / kScanBlock.  `offsets` may be null to compute
// the total only.  Defined in marching_cubes.hip.
int launch_exclusive_scan(const int32_t* counts, int n, long long* offsets, long long* sums,
                          hipStream_t s);

// One workgroup: tile_off[i] = sum of tile_sum[0..i) for i <= n (int64;
// tile_off[n] is the total); with tile_nne, also group_off[i] = sum of
// ceil(tile_nne / 4) over the tiles before i (group_off[n] the total).  The
// arrays must be 16-byte aligned.  Defined in marching_cubes.hip.
int launch_tile_scan(const int32_t* tile_sum, int n, long long* tile_off, hipStream_t s,
                     const int32_t* tile_nne = nullptr, int32_t* group_off = nullptr);

}  // namespace recon
