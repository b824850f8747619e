// scan.hpp -- device-wide exclusive scan of int32 counts into int64 offsets
// (three launches: block sums, one-workgroup scan of the block sums, apply),
// shared by marching cubes and the differentiable path.
#pragma once

#include "recon_common.hpp"

namespace recon {

// ---- exclusive scan of int32 counts into int64 offsets ---------------------
constexpr int kScanItems = 4;  // scan_apply_kernel's vector path assumes 4
constexpr int kScanBlock = 256 * kScanItems;

// A 64-bit DPP move: the same move applied to both halves.
template <int Ctrl, int RowMask, int BankMask, bool BoundCtrl>
__device__ __forceinline__ long long dpp64(long long v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)v, Ctrl, RowMask, BankMask, BoundCtrl);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), Ctrl, RowMask, BankMask, BoundCtrl);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo);
}

// Inclusive prefix sum of a 64-bit value over the 64 lanes of a wavefront:
// the moves of wave_inclusive_scan on both halves, 64-bit adds.
__device__ __forceinline__ long long wave_inclusive_scan64(long long v) {
    long long t = v;
    t += dpp64<0x111, 0xF, 0xF, true>(v);   // row_shr:1
    t += dpp64<0x112, 0xF, 0xF, true>(v);   // row_shr:2
    t += dpp64<0x113, 0xF, 0xF, true>(v);   // row_shr:3
    t += dpp64<0x114, 0xF, 0xE, true>(t);   // row_shr:4, banks 1-3
    t += dpp64<0x118, 0xF, 0xC, true>(t);   // row_shr:8, banks 2-3
    t += dpp64<0x142, 0xA, 0xF, false>(t);  // row_bcast:15, rows 1, 3
    t += dpp64<0x143, 0xC, 0xF, false>(t);  // row_bcast:31, rows 2, 3
    return t;
}

// Exclusive prefix over a 256-thread workgroup of one 64-bit value per
// thread (wave scans by DPP, the four wave totals through LDS); *total is
// the workgroup's sum.  sh: 4 or more long longs of LDS; every thread calls.
__device__ __forceinline__ long long block_exclusive_scan(long long v, long long* sh,
                                                          long long* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long incl = wave_inclusive_scan64(v);
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    long long before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        before += w < wave ? sh[w] : 0;
        all += sh[w];
    }
    *total = all;
    __syncthreads();  // sh is reused by the next call
    return before + incl - v;
}

// Exclusive scan of n int32 counts into int64 offsets on stream s: three
// launches (block sums, top-level scan, apply); sums[nb] receives the total,
// nb = (n + kScanBlock - 1) / kScanBlock.  `offsets` may be null to compute
// the total only.  Defined in marching_cubes.hip.
int launch_exclusive_scan(const int32_t* counts, int n, long long* offsets, long long* sums,
                          hipStream_t s);

// Arrays of equal length scanned by one set of the three launches (the
// launches' blockIdx.y picks the array): in[i] into out[i], with sums[i]
// as above (nb + 1 long longs each); the apply launch runs when every out[i]
// is set.
struct ScanJobs {
    const int32_t* in[2];
    long long* sums[2];
    long long* out[2];
};
int launch_exclusive_scans(const ScanJobs& jobs, int count, int n, hipStream_t s);

// One workgroup: tile_off[i] = sum of tile_sum[0..i) for i <= n (int64;
// tile_off[n] is the total); with tile_nne, also group_off[i] = sum of
// ceil(tile_nne / 4) over the tiles before i (group_off[n] the total).  The
// arrays must be 16-byte aligned.  Defined in marching_cubes.hip.
int launch_tile_scan(const int32_t* tile_sum, int n, long long* tile_off, hipStream_t s,
                     const int32_t* tile_nne = nullptr, int32_t* group_off = nullptr);

}  // namespace recon
