// diff_host.hpp -- host-side pieces of the differentiable path shared by the
// single-volume entry points (tsdf_diff.hip) and the batched refinement
// (diff_batch.hip).  Defined in tsdf_diff.hip.
#pragma once

#include "recon_common.hpp"

namespace recon {

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Scratch of the pair merge for up to n pairs (byte offsets from its base).
struct MergeLayout {
    size_t head, offsets, sums, uid, lam, acc, maxbits, block_max, total;
    int nb;
};
MergeLayout merge_layout(int64_t n);

// The merged list of n (voxel, value) pairs (voxel indices below N < 2^32,
// n < 2^31): one pair per voxel in the order of first occurrence, values
// summed exactly (int64 fixed point, order-independent).  `owner` holds N
// uint32, all ones before and after.  On return *uid / *lam point into the
// scratch and *count at the device count of merged pairs.
int launch_merge(const int64_t* ids, const float* lam, int64_t n, int64_t N, uint32_t* owner,
                 char* scratch, const MergeLayout& M, uint32_t** uid, float** lam_u,
                 const long long** count, hipStream_t s);

// out[i] = scale * (sum of in[i*n .. i*n + n-1]) for i < nout, in a fixed
// order (one workgroup per output).
int launch_fixed_sum(const double* in, int64_t n, double* out, int nout, double scale,
                     hipStream_t s);

int launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t s);

}  // namespace recon
