// tsdf_diff.hip -- the differentiable integrate, the surface loss and their
// adjoint for gfx950 (an extension of cu-Reconstruction's TSDF fusion path;
// the reference has no counterpart).  Device code in diff_device.hpp.
//
// Poses are device-resident float32 world->camera matrices, 12 per frame
// (row-major R, then t), so that a caller can update them on the device
// between steps.  Each call builds its frame table on the device
// (diff_setup_kernel), so none reads host memory after it is queued.
//
//   recon_tsdf_diff_integrate  per chunk of <= 64 frames: depth bounds, the
//                              conservative cull of the plain path (a
//                              bilinear sample is a convex combination of
//                              taps inside the plain path's footprint hull,
//                              so its bounds still hold), then a persistent
//                              list kernel doing bilinear updates;
//   recon_tsdf_surface_loss    one thread per point: trilinear residual,
//                              per-block float64 loss partials reduced in a
//                              fixed order, and the 8 (voxel, dL/dD) pairs of
//                              the point;
//   recon_tsdf_diff_adjoint    (voxel, dL/dD) pairs over a bounded grid.
//                              Up to 64 frames: lanes are frames, 64 / F
//                              voxels per wave, each (voxel, frame) evaluated
//                              once; the updates before a frame come from a
//                              ballot, the chain of running averages from a
//                              suffix-product scan, and each lane keeps its
//                              frame's pose sums in float64 registers.
//                              Longer sequences: one voxel per lane, frames
//                              replayed forward (count) then backward.  Per
//                              wave partials are summed in a fixed order
//                              (run-to-run identical); depth gradients by
//                              float atomics.
#pragma clang fp contract(off)

#include "diff_device.hpp"
#include "diff_host.hpp"
#include "scan.hpp"

#include <algorithm>
#include <cstring>

namespace recon {

using Shape = DefaultShape;
constexpr int kDiffChunk = 64;          // frames per cull mask
constexpr int kAdjointMaxWaves = 8192;  // bound on the adjoint grid

__global__ __launch_bounds__(64) void diff_setup_kernel(const VolumeParams vp,
                                                        const float* __restrict__ poses,
                                                        const int nframes, const float* depth0,
                                                        const size_t frame_stride, const int pitch,
                                                        float* bounds, const size_t bounds_per,
                                                        FrameParams* __restrict__ table) {
    const int f = blockIdx.x * 64 + threadIdx.x;
    if (f >= nframes) return;
    FrameParams fp;
    for (int i = 0; i < 9; ++i) fp.R[i] = poses[12 * f + i];
    for (int i = 0; i < 3; ++i) fp.t[i] = poses[12 * f + 9 + i];
    fp.depth = depth0 + frame_stride * f;
    fp.bound0 = bounds + bounds_per * (f % kDiffChunk);
    fp.bound1 = fp.bound0 + (size_t)vp.b0x * vp.b0y;
    fp.pitch = pitch;
    frame_planes(vp, &fp);
    table[f] = fp;
}

__global__ __launch_bounds__(256) void diff_bounds_kernel(const BoundsJob job, size_t frame_stride,
                                                          size_t bounds_stride,
                                                          unsigned* list_count) {
    __shared__ int s0[16];
    __shared__ int s1[4];
    if (threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) *list_count = 0;
    BoundsJob j = job;
    j.depth += frame_stride * blockIdx.z;
    j.bound0 += bounds_stride * blockIdx.z;
    j.bound1 += bounds_stride * blockIdx.z;
    bounds_block(j, blockIdx.x, blockIdx.y, s0, s1);
}

// 64 tiles per workgroup of W = blockDim.x / 64 wavefronts (at most
// kCullWaves); wave g tests frames g, g + W, ... of every tile, and wave 0
// ORs the tile's masks and appends it.
__global__ __launch_bounds__(64 * kCullWaves) void diff_cull_kernel(
    const VolumeParams vp, const int nframes, const FrameTable ft, ListEntry* __restrict__ list,
    unsigned* __restrict__ list_count) {
    __shared__ uint64_t s_live[kCullWaves][64];
    const uint32_t ntiles = (uint32_t)vp.tiles_x * vp.tiles_y * vp.tiles_z;
    const int i = threadIdx.x & 63, g = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    const uint32_t id = blockIdx.x * 64u + i;
    s_live[g][i] = id < ntiles ? tile_frames<Shape>(vp, nframes, ft, id, g, nw) : 0ull;
    __syncthreads();
    if (g != 0) return;
    uint64_t live = 0;
    for (int k = 0; k < nw; ++k) live |= s_live[k][i];
    append_wave(live, id, 0u, list, list_count);
}

template <bool VEC>
__global__ __launch_bounds__(kThreads) void diff_integrate_kernel(
    const VolumeParams vp, const int nframes, const FrameTable ft,
    const ListEntry* __restrict__ list, const unsigned* __restrict__ list_count) {
    __shared__ float rtab[kRcpTable];
    rtab[threadIdx.x] = 1.0f / (float)(threadIdx.x + 1);
    __syncthreads();
    const unsigned n = __builtin_amdgcn_readfirstlane(*list_count);
    const unsigned waves = gridDim.x * (kThreads / 64);
    for (unsigned i = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t tile = __builtin_amdgcn_readfirstlane(list[i].tile);
        const uint64_t frames =
            ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(list[i].frames >> 32)) << 32) |
            __builtin_amdgcn_readfirstlane((uint32_t)list[i].frames);
        int x0, y0, z0;
        tile_origin<Shape>(vp, tile, x0, y0, z0);
        diff_integrate_tile<Shape, FrameTable, VEC>(vp, nframes, ft, frames, x0, y0, z0, rtab);
    }
}

// ---- surface loss ----------------------------------------------------------------
__global__ __launch_bounds__(256) void surface_loss_kernel(
    const VolumeParams vp, const float* __restrict__ points, const float* __restrict__ weights,
    const int64_t npoints, float* __restrict__ residuals, double* __restrict__ partials,
    int64_t* __restrict__ lam_ids, float* __restrict__ lam) {
    __shared__ double s_wave[4];
    const int64_t k = blockIdx.x * 256ll + threadIdx.x;
    double l = 0.0;
    if (k < npoints) {
        const float p[3] = {points[3 * k], points[3 * k + 1], points[3 * k + 2]};
        const float w = weights ? weights[k] : 1.0f;
        int64_t ids[8];
        float coef[8];
        const bool ok = tri_support(vp, p, ids, coef);
        float r = 0.0f;
        if (ok) {
#pragma unroll
            for (int i = 0; i < 8; ++i) r = r + coef[i] * vp.D[ids[i]];
        }
        if (residuals) residuals[k] = r;
        l = ok ? (double)w * (double)r * (double)r : 0.0;
        if (lam_ids) {
            // dL/dD of each support voxel: 2 w r coef / K
            const float g = (2.0f * w * r) / (float)npoints;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                lam_ids[8 * k + i] = ok ? ids[i] : -1;
                lam[8 * k + i] = ok ? g * coef[i] : 0.0f;
            }
        }
    }
    l = wave_sum_f64(l);
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = ((s_wave[0] + s_wave[1]) + s_wave[2]) + s_wave[3];
}

// Sums `n` values per output (stride `n`, `nout` outputs, one block each) in
// a fixed order: thread t takes t, t+256, ..., then a fixed tree.
__global__ __launch_bounds__(256) void fixed_sum_kernel(const double* __restrict__ in,
                                                        const int64_t n, double* __restrict__ out,
                                                        const double scale) {
    __shared__ double s[256];
    const double* src = in + (size_t)blockIdx.x * n;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) acc += src[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int m = 128; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) s[threadIdx.x] += s[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0] * scale;
}

// ---- adjoint ---------------------------------------------------------------------
// Sequences of up to 64 frames: lanes are frames.  A wave holds S = 64 / F
// segments of F lanes; segment s takes entry e = (wave + k * waves) * S + s
// and its lane f evaluates frame f of that voxel once.  The updates before
// frame f are the set bits of the segment's ballot below the lane; dL/dD
// after frame f is lambda times the product of a_g over the segment's later
// frames (a suffix-product scan); each lane accumulates its frame's six pose
// sums over all the wave's entries.  At the end the segments are summed in
// order into partial[(f*6 + c) * waves + wave].
template <class IdT>
__global__ __launch_bounds__(256) void diff_adjoint_kernel(
    const VolumeParams vp, const int nframes, const FrameTable ft, const IdT* __restrict__ ids,
    const float* __restrict__ lam_in, const int64_t n_host, const long long* __restrict__ n_dev,
    double* __restrict__ partial, unsigned long long* __restrict__ gdepth,
    const size_t gd_frame_stride, const unsigned* __restrict__ gd_maxbits) {
    const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
    const unsigned waves = gridDim.x * 4u;
    const unsigned wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const LaneLayout ly = lane_layout(nframes);
    const FrameParams L = ft.f[ly.ok ? ly.f : 0];  // kept in registers
    unsigned long long* gd_frame = gdepth ? gdepth + gd_frame_stride * (ly.ok ? ly.f : 0) : nullptr;
    const double gd_scale = gdepth ? depth_grad_scale(*gd_maxbits) : 1.0;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t e0 = (int64_t)wave * ly.S; e0 < n; e0 += (int64_t)waves * ly.S) {
        const int64_t e = e0 + ly.seg;
        int64_t v = -1;
        float lam = 0.0f;
        if (ly.ok && e < n) {
            v = (int64_t)ids[e];
            lam = lam_in[e];
        }
        adjoint_lane_pass(vp, ly, L, v, lam, gd_frame, gd_scale, acc);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const double sum = segment_sum(acc[c], ly);
        if (ly.seg == 0 && ly.ok) partial[((size_t)ly.f * 6 + c) * waves + wave] = sum;
    }
}

// Longer sequences: one voxel per lane, frames in a loop (count, then
// reverse).  Entry e = wave * 64 + lane + k * (waves * 64);
// partial[(f*6 + c) * waves + wave] is owned by one wave: written on its
// first pass, then accumulated.
template <class IdT>
__global__ __launch_bounds__(256) void diff_adjoint_long_kernel(
    const VolumeParams vp, const int nframes, const FrameTable ft, const IdT* __restrict__ ids,
    const float* __restrict__ lam_in, const int64_t n_host, const long long* __restrict__ n_dev,
    double* __restrict__ partial, unsigned long long* __restrict__ gdepth,
    const size_t gd_frame_stride, const unsigned* __restrict__ gd_maxbits) {
    const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
    const int lane = threadIdx.x & 63;
    const double gd_scale = gdepth ? depth_grad_scale(*gd_maxbits) : 1.0;
    const unsigned waves = gridDim.x * 4u;
    const unsigned wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int64_t N = (int64_t)vp.gx * vp.gy * vp.gz;
    bool first = true;
    for (int64_t e0 = (int64_t)wave * 64; e0 < n; e0 += (int64_t)waves * 64) {
        const int64_t e = e0 + lane;
        int64_t v = -1;
        float lam = 0.0f;
        if (e < n) {
            v = (int64_t)ids[e];
            lam = lam_in[e];
        }
        const bool act = v >= 0 && v < N && lam != 0.0f && isfinite(lam);
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        if (act) {
            const int64_t x = v % vp.gx, yz = v / vp.gx;
            cx = voxel_centre((int)x, vp.vs[0], vp.off[0]);
            cy = voxel_centre((int)(yz % vp.gy), vp.vs[1], vp.off[1]);
            cz = voxel_centre((int)(yz / vp.gy), vp.vs[2], vp.off[2]);
        }
        // the voxel's updates, as the forward made them
        int cnt = 0;
        if (act) {
            for (int f = 0; f < nframes; ++f) {
                const FrameParams& F = ft.f[f];
                const BiProj s = bi_project(vp, F, cx, cy, cz);
                if (!s.in) continue;
                const float* q = F.depth + s.addr;
                const float t00 = q[0], t10 = q[1], t01 = q[F.pitch], t11 = q[F.pitch + 1];
                const bool valid = tap_valid(t00, vp.maxd) && tap_valid(t10, vp.maxd) &&
                                   tap_valid(t01, vp.maxd) && tap_valid(t11, vp.maxd);
                const float sdf = bi_depth(s.a, s.b, t00, t10, t01, t11) - s.pz;
                cnt += (valid && sdf >= -vp.trunc) ? 1 : 0;
            }
        }
        float lc = lam;
        for (int f = nframes - 1; f >= 0; --f) {
            double c[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (act && cnt > 0) {
                const FrameParams& F = ft.f[f];
                const BiProj s = bi_project(vp, F, cx, cy, cz);
                if (s.in) {
                    const float* q = F.depth + s.addr;
                    const float t00 = q[0], t10 = q[1], t01 = q[F.pitch], t11 = q[F.pitch + 1];
                    const bool valid = tap_valid(t00, vp.maxd) && tap_valid(t10, vp.maxd) &&
                                       tap_valid(t01, vp.maxd) && tap_valid(t11, vp.maxd);
                    const float sdf = bi_depth(s.a, s.b, t00, t10, t01, t11) - s.pz;
                    if (valid && sdf >= -vp.trunc) {
                        --cnt;
                        const float wb = fminf((float)cnt, vp.maxw);
                        const float a = wb / (wb + 1.0f), b = 1.0f / (wb + 1.0f);
                        const float g = (sdf > 0.0f && sdf >= vp.trunc) ? 0.0f : lc * b;
                        if (g != 0.0f) {
                            const float ia = 1.0f - s.a, ib = 1.0f - s.b;
                            const float dd_du = ib * (t10 - t00) + s.b * (t11 - t01);
                            const float dd_dv = ia * (t01 - t00) + s.a * (t11 - t10);
                            const float iz = 1.0f / s.pz;
                            const float gx = ((g * dd_du) * vp.fx) * iz;
                            const float gy = ((g * dd_dv) * vp.fy) * iz;
                            const float gz = -((gx * s.px + gy * s.py) * iz) - g;
                            c[0] = gx;
                            c[1] = gy;
                            c[2] = gz;
                            c[3] = (double)s.py * gz - (double)s.pz * gy;
                            c[4] = (double)s.pz * gx - (double)s.px * gz;
                            c[5] = (double)s.px * gy - (double)s.py * gx;
                            if (gdepth) {
                                unsigned long long* o = gdepth + gd_frame_stride * f + s.addr;
                                depth_grad_add(o, g * (ia * ib), gd_scale);
                                depth_grad_add(o + 1, g * (s.a * ib), gd_scale);
                                depth_grad_add(o + F.pitch, g * (ia * s.b), gd_scale);
                                depth_grad_add(o + F.pitch + 1, g * (s.a * s.b), gd_scale);
                            }
                        }
                        lc = lc * a;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) c[k] = wave_sum_f64(c[k]);
            if (lane == 0) {
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    double* slot = partial + ((size_t)f * 6 + k) * waves + wave;
                    *slot = first ? c[k] : *slot + c[k];
                }
            }
        }
        first = false;
    }
    // a wave without entries still owns its slots
    if (first && lane == 0)
        for (int f = 0; f < nframes; ++f)
            for (int k = 0; k < 6; ++k) partial[((size_t)f * 6 + k) * waves + wave] = 0.0;
}

// ---- merging repeated voxels -------------------------------------------------------
// (voxel, value) pairs -> one pair per voxel, in the order of first
// occurrence, with the values summed exactly: every value is scaled to an
// int64 fixed-point number (the scale puts the largest |value| at
// 2^(61 - ceil(log2(n+1))), so no sum of n terms overflows) and added with
// integer atomics, whose result does not depend on their order.  `owner`
// (one uint32 per voxel, all ones between calls) takes the smallest pair
// index of each voxel and is restored at the end.
__device__ __forceinline__ bool pair_valid(int64_t v, float l, int64_t N) {
    return v >= 0 && v < N && l != 0.0f && isfinite(l);
}

__device__ __forceinline__ double merge_scale(unsigned maxbits, int lg) {
    const float m = __uint_as_float(maxbits);
    return m > 0.0f ? ldexp(1.0, 61 - lg - ilogbf(m) - 1) : 1.0;
}

__global__ __launch_bounds__(256) void merge_mark_kernel(const int64_t* __restrict__ ids,
                                                         const float* __restrict__ lam, int64_t n,
                                                         int64_t N, uint32_t* __restrict__ owner,
                                                         unsigned* __restrict__ block_max) {
    __shared__ unsigned s_max[4];
    const int64_t e = blockIdx.x * 256ll + threadIdx.x;
    unsigned m = 0;
    if (e < n) {
        const int64_t v = ids[e];
        const float l = lam[e];
        if (pair_valid(v, l, N)) {
            // the plain read skips pairs an earlier one already beat
            if ((uint32_t)e < owner[v]) atomicMin(owner + v, (uint32_t)e);
            m = __float_as_uint(fabsf(l));  // orders like the magnitudes
        }
    }
    // the block's largest magnitude into a slot of its own (atomics on one
    // address from every wave would serialise)
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, d, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) block_max[blockIdx.x] = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
}

// One workgroup: the largest of the block maxima.
__global__ __launch_bounds__(256) void merge_max_kernel(const unsigned* __restrict__ block_max,
                                                        int nblocks, unsigned* __restrict__ maxbits) {
    __shared__ unsigned s[256];
    unsigned m = 0;
    for (int i = threadIdx.x; i < nblocks; i += 256) m = max(m, block_max[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] = max(s[threadIdx.x], s[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *maxbits = s[0];
}

__global__ __launch_bounds__(256) void merge_flag_kernel(const int64_t* __restrict__ ids,
                                                         const float* __restrict__ lam, int64_t n,
                                                         int64_t N, const uint32_t* __restrict__ owner,
                                                         int32_t* __restrict__ head) {
    const int64_t e = blockIdx.x * 256ll + threadIdx.x;
    if (e >= n) return;
    const int64_t v = ids[e];
    head[e] = pair_valid(v, lam[e], N) && owner[v] == (uint32_t)e ? 1 : 0;
}

__global__ __launch_bounds__(256) void merge_slot_kernel(const int64_t* __restrict__ ids, int64_t n,
                                                         const int32_t* __restrict__ head,
                                                         const long long* __restrict__ offsets,
                                                         uint32_t* __restrict__ uid,
                                                         long long* __restrict__ acc) {
    const int64_t e = blockIdx.x * 256ll + threadIdx.x;
    if (e >= n || !head[e]) return;
    const long long slot = offsets[e];
    uid[slot] = (uint32_t)ids[e];
    acc[slot] = 0;
}

__global__ __launch_bounds__(256) void merge_add_kernel(const int64_t* __restrict__ ids,
                                                        const float* __restrict__ lam, int64_t n,
                                                        int64_t N, const uint32_t* __restrict__ owner,
                                                        const long long* __restrict__ offsets,
                                                        const unsigned* __restrict__ maxbits, int lg,
                                                        long long* __restrict__ acc) {
    const int64_t e = blockIdx.x * 256ll + threadIdx.x;
    if (e >= n) return;
    const int64_t v = ids[e];
    const float l = lam[e];
    if (!pair_valid(v, l, N)) return;
    const double scale = merge_scale(*maxbits, lg);
    const long long q = llrint((double)l * scale);
    atomicAdd(reinterpret_cast<unsigned long long*>(acc + offsets[owner[v]]),
              (unsigned long long)q);
}

__global__ __launch_bounds__(256) void merge_finish_kernel(const long long* __restrict__ count,
                                                           const uint32_t* __restrict__ uid,
                                                           const long long* __restrict__ acc,
                                                           const unsigned* __restrict__ maxbits,
                                                           int lg, float* __restrict__ lam_u,
                                                           uint32_t* __restrict__ owner) {
    const int64_t i = blockIdx.x * 256ll + threadIdx.x;
    if (i >= *count) return;
    lam_u[i] = (float)((double)acc[i] / merge_scale(*maxbits, lg));
    owner[uid[i]] = 0xFFFFFFFFu;
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* __restrict__ p, int64_t n,
                                                       uint32_t value) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = value;
}

// Block maxima of |lam| over n values (float bits, which order like the
// magnitudes), for the depth-gradient scale; n may live on the device.
__global__ __launch_bounds__(256) void lambda_max_kernel(const float* __restrict__ lam,
                                                         const int64_t n_host,
                                                         const long long* __restrict__ n_dev,
                                                         unsigned* __restrict__ block_max) {
    __shared__ unsigned s_max[4];
    const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
    unsigned m = 0;
    for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const float l = lam[e];
        if (isfinite(l)) m = max(m, __float_as_uint(fabsf(l)));
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, d, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) block_max[blockIdx.x] = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
}

__global__ __launch_bounds__(256) void zero_u64_kernel(unsigned long long* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) p[i] = 0ull;
}

// gd (frame y at y * frame_stride floats) <- the accumulators (frame y at
// y * per_frame) divided by their scale.
__global__ __launch_bounds__(256) void depth_grad_finish_kernel(
    const unsigned long long* __restrict__ acc, size_t per_frame, float* __restrict__ gd,
    size_t frame_stride, const unsigned* __restrict__ maxbits) {
    const double inv = 1.0 / depth_grad_scale(*maxbits);
    const unsigned long long* a = acc + per_frame * blockIdx.y;
    float* out = gd + frame_stride * blockIdx.y;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < per_frame; i += gridDim.x * 256ull)
        out[i] = (float)((double)(long long)a[i] * inv);
}

__global__ __launch_bounds__(256) void zero_frames_kernel(float* __restrict__ p, size_t frame_stride,
                                                          size_t per_frame) {
    float* base = p + frame_stride * blockIdx.y;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < per_frame; i += gridDim.x * 256ull)
        base[i] = 0.0f;
}

// ---- host side ----------------------------------------------------------------------

// Grows the handle's scratch for the differentiable path to `bytes`.  A
// buffer being replaced may still be read by queued kernels, so the stream is
// drained first.
static int ensure_diff_scratch(recon_tsdf* h, size_t bytes, hipStream_t s) {
    if (h->diff_scratch_bytes >= bytes) return RECON_OK;
    if (h->diff_scratch) {
        RECON_HIP_TRY(hipStreamSynchronize(s), "diff scratch resize synchronize");
        (void)hipFree(h->diff_scratch);
    }
    h->diff_scratch = nullptr;
    h->diff_scratch_bytes = 0;
    RECON_HIP_TRY(hipMalloc(&h->diff_scratch, bytes), "diff scratch");
    h->diff_scratch_bytes = bytes;
    return RECON_OK;
}

static int ensure_list(recon_tsdf* h, const VolumeParams& vp) {
    return ensure_cull_list(h, (size_t)vp.tiles_x * vp.tiles_y * vp.tiles_z, sizeof(ListEntry));
}

static VolumeParams diff_volume_params(const recon_tsdf* h, int cols, int rows) {
    VolumeParams vp = h->vp;
    vp.cols = cols;
    vp.rows = rows;
    vp.fcols = (float)cols;
    vp.frows = (float)rows;
    vp.b0x = (cols + kBound0 - 1) / kBound0;
    vp.b0y = (rows + kBound0 - 1) / kBound0;
    vp.b1x = (cols + kBound1 - 1) / kBound1;
    vp.b1y = (rows + kBound1 - 1) / kBound1;
    return vp;
}

static int check_frames(const recon_tsdf* h, const recon_depth_view* depth0,
                        size_t frame_stride_bytes, int32_t num_frames, const float* poses,
                        const char* fn) {
    if (!h) return fail(RECON_ERR_NULL, "%s: handle is NULL", fn);
    if (h->info.weight_type != RECON_WEIGHT_BY_DEPTH)
        return fail(RECON_ERR_CONFIG, "%s: the differentiable integrate weighs every observation 1 "
                    "(WEIGHT_BY_DEPTH)", fn);
    if (!depth0 || !depth0->data) return fail(RECON_ERR_NULL, "%s: depth is NULL", fn);
    if (!poses && num_frames > 0) return fail(RECON_ERR_NULL, "%s: poses is NULL", fn);
    if (num_frames < 0) return fail(RECON_ERR_SIZE, "%s: num_frames < 0", fn);
    if (depth0->rows < 2 || depth0->cols < 2)
        return fail(RECON_ERR_SIZE, "%s: depth %dx%d must be at least 2x2 for bilinear taps", fn,
                    depth0->cols, depth0->rows);
    if (depth0->step < 4 * (size_t)depth0->cols || depth0->step % 4 != 0)
        return fail(RECON_ERR_SIZE, "%s: depth step %zu must be a multiple of 4 and >= 4*cols", fn,
                    depth0->step);
    if ((depth0->step / 4) * (size_t)depth0->rows >= (size_t(1) << 31))
        return fail(RECON_ERR_RANGE, "%s: depth image exceeds 2^31 floats", fn);
    if (num_frames > 1 && (frame_stride_bytes % 4 != 0 ||
                           frame_stride_bytes < depth0->step * (size_t)depth0->rows))
        return fail(RECON_ERR_SIZE, "%s: frame stride %zu must be a multiple of 4 and hold a frame",
                    fn, frame_stride_bytes);
    return RECON_OK;
}

struct DiffLayout {
    size_t table, bounds, partial, total;
};

static DiffLayout diff_layout(const VolumeParams& vp, int nframes, size_t partial_doubles) {
    DiffLayout L;
    const size_t per = (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
    L.table = 0;
    L.bounds = align256(sizeof(FrameParams) * (size_t)std::max(nframes, 1));
    L.partial = L.bounds + align256(per * kDiffChunk * sizeof(float));
    L.total = L.partial + align256(partial_doubles * sizeof(double));
    return L;
}

MergeLayout merge_layout(int64_t n) {
    MergeLayout M;
    const size_t un = (size_t)std::max<int64_t>(n, 1);
    M.nb = (int)((un + kScanBlock - 1) / kScanBlock);
    M.head = 0;
    M.offsets = M.head + align256(4 * un);
    M.sums = M.offsets + align256(8 * un);
    M.uid = M.sums + align256(8 * ((size_t)M.nb + 1));
    M.lam = M.uid + align256(4 * un);
    M.acc = M.lam + align256(4 * un);
    M.maxbits = M.acc + align256(8 * un);
    M.block_max = M.maxbits + 256;
    M.total = M.block_max + align256(4 * ((un + 255) / 256));
    return M;
}

int launch_merge(const int64_t* ids, const float* lam, int64_t n, int64_t N, uint32_t* owner,
                 char* mb, const MergeLayout& M, uint32_t** uid_out, float** lam_out,
                 const long long** count_out, hipStream_t s) {
    int32_t* head = reinterpret_cast<int32_t*>(mb + M.head);
    long long* offsets = reinterpret_cast<long long*>(mb + M.offsets);
    long long* sums = reinterpret_cast<long long*>(mb + M.sums);
    uint32_t* uid = reinterpret_cast<uint32_t*>(mb + M.uid);
    float* lam_u = reinterpret_cast<float*>(mb + M.lam);
    long long* acc = reinterpret_cast<long long*>(mb + M.acc);
    unsigned* maxbits = reinterpret_cast<unsigned*>(mb + M.maxbits);
    unsigned* block_max = reinterpret_cast<unsigned*>(mb + M.block_max);
    const int nblocks = (int)((n + 255) / 256);
    const dim3 pg((unsigned)nblocks);
    const int lg = 64 - __builtin_clzll((unsigned long long)n);
    hipLaunchKernelGGL(merge_mark_kernel, pg, dim3(256), 0, s, ids, lam, n, N, owner, block_max);
    hipLaunchKernelGGL(merge_max_kernel, dim3(1), dim3(256), 0, s, block_max, nblocks, maxbits);
    hipLaunchKernelGGL(merge_flag_kernel, pg, dim3(256), 0, s, ids, lam, n, N, owner, head);
    RECON_HIP_TRY(hipGetLastError(), "merge kernels launch");
    int rc = launch_exclusive_scan(head, (int)n, offsets, sums, s);
    if (rc) return rc;
    const long long* count = sums + M.nb;
    hipLaunchKernelGGL(merge_slot_kernel, pg, dim3(256), 0, s, ids, n, head, offsets, uid, acc);
    hipLaunchKernelGGL(merge_add_kernel, pg, dim3(256), 0, s, ids, lam, n, N, owner, offsets,
                       maxbits, lg, acc);
    hipLaunchKernelGGL(merge_finish_kernel, pg, dim3(256), 0, s, count, uid, acc, maxbits, lg,
                       lam_u, owner);
    RECON_HIP_TRY(hipGetLastError(), "merge kernels launch");
    *uid_out = uid;
    *lam_out = lam_u;
    *count_out = count;
    return RECON_OK;
}

int launch_fixed_sum(const double* in, int64_t n, double* out, int nout, double scale,
                     hipStream_t s) {
    if (nout <= 0) return RECON_OK;
    hipLaunchKernelGGL(fixed_sum_kernel, dim3(nout), dim3(256), 0, s, in, n, out, scale);
    RECON_HIP_TRY(hipGetLastError(), "fixed_sum_kernel launch");
    return RECON_OK;
}

int launch_fill_u32(uint32_t* p, int64_t n, uint32_t value, hipStream_t s) {
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, p, n, value);
    RECON_HIP_TRY(hipGetLastError(), "fill_u32_kernel launch");
    return RECON_OK;
}

// The per-voxel owner table of the merge (all ones between calls), made on
// first use.
static int ensure_owner(recon_tsdf* h, hipStream_t s) {
    if (h->owner && !h->owner_dirty) return RECON_OK;
    const int64_t n = h->info.num_voxels;
    if (!h->owner)
        RECON_HIP_TRY(hipMalloc((void**)&h->owner, (size_t)n * sizeof(uint32_t)),
                      "merge owner table");
    h->owner_dirty = false;
    return launch_fill_u32(h->owner, n, 0xFFFFFFFFu, s);
}

// Frame table of all frames into the scratch (device side, from the poses).
static void launch_setup(const VolumeParams& vp, const recon_depth_view* depth0,
                         size_t frame_stride_bytes, int nframes, const float* poses, char* base,
                         const DiffLayout& L, hipStream_t s) {
    const size_t per = (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
    hipLaunchKernelGGL(diff_setup_kernel, dim3((nframes + 63) / 64), dim3(64), 0, s, vp, poses,
                       nframes, depth0->data, frame_stride_bytes / 4, (int)(depth0->step / 4),
                       reinterpret_cast<float*>(base + L.bounds), per,
                       reinterpret_cast<FrameParams*>(base + L.table));
}

static int diff_occupancy() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, diff_integrate_kernel<true>, kThreads,
                                                     0) != hipSuccess ||
        blocks < 1)
        blocks = 2;
    return blocks;
}

}  // namespace recon

using namespace recon;

extern "C" int recon_pose_to_matrix(const double* poses, int32_t num_poses, float* rt) {
    clear_error();
    if ((!poses || !rt) && num_poses > 0) return fail(RECON_ERR_NULL, "recon_pose_to_matrix: NULL");
    if (num_poses < 0) return fail(RECON_ERR_SIZE, "recon_pose_to_matrix: num_poses < 0");
    for (int i = 0; i < num_poses; ++i) {
        if (!pose_is_finite(poses + 7 * i))
            return fail(RECON_ERR_CONFIG, "recon_pose_to_matrix: pose %d non-finite", i);
        pose_to_matrix_f32(poses + 7 * i, rt + 12 * i, rt + 12 * i + 9);
    }
    return RECON_OK;
}

extern "C" int recon_tsdf_diff_integrate(recon_tsdf* h, const recon_depth_view* depth0,
                                         size_t frame_stride_bytes, int32_t num_frames,
                                         const float* poses, recon_stream_t stream) {
    clear_error();
    int rc = check_frames(h, depth0, frame_stride_bytes, num_frames, poses,
                          "recon_tsdf_diff_integrate");
    if (rc) return rc;
    if (num_frames == 0) return RECON_OK;
    rc = require_gfx950();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const VolumeParams vp = diff_volume_params(h, depth0->cols, depth0->rows);
    const DiffLayout L = diff_layout(vp, num_frames, 0);
    if ((rc = ensure_diff_scratch(h, L.total, s))) return rc;
    if ((rc = ensure_list(h, vp))) return rc;
    char* base = static_cast<char*>(h->diff_scratch);
    launch_setup(vp, depth0, frame_stride_bytes, num_frames, poses, base, L, s);
    RECON_HIP_TRY(hipGetLastError(), "diff_setup_kernel launch");
    const FrameParams* table = reinterpret_cast<const FrameParams*>(base + L.table);
    const size_t per = (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
    const int ntiles = vp.tiles_x * vp.tiles_y * vp.tiles_z;
    ListEntry* list = static_cast<ListEntry*>(h->list);
    static const int occ = diff_occupancy();
    const dim3 grid(h->num_cus * occ);
    for (int f0 = 0; f0 < num_frames; f0 += kDiffChunk) {
        const int n = std::min(kDiffChunk, num_frames - f0);
        BoundsJob job;
        job.depth = reinterpret_cast<const float*>(reinterpret_cast<const char*>(depth0->data) +
                                                   frame_stride_bytes * (size_t)f0);
        job.bound0 = reinterpret_cast<float*>(base + L.bounds);
        job.bound1 = job.bound0 + (size_t)vp.b0x * vp.b0y;
        job.pitch = (int32_t)(depth0->step / 4);
        job.cols = vp.cols;
        job.rows = vp.rows;
        job.b0x = vp.b0x;
        job.b0y = vp.b0y;
        job.b1x = vp.b1x;
        job.b1y = vp.b1y;
        job.maxd = vp.maxd;
        hipLaunchKernelGGL(diff_bounds_kernel, dim3(vp.b1x, vp.b1y, n), dim3(256), 0, s, job,
                           frame_stride_bytes / 4, per, h->list_count);
        RECON_HIP_TRY(hipGetLastError(), "diff_bounds_kernel launch");
        const FrameTable ft{table + f0};
        const unsigned groups = (unsigned)((ntiles + 63) / 64);
        hipLaunchKernelGGL(diff_cull_kernel, dim3(groups),
                           dim3(64 * cull_waves(groups, n, h->num_cus)), 0, s, vp, n, ft, list,
                           h->list_count);
        RECON_HIP_TRY(hipGetLastError(), "diff_cull_kernel launch");
        if (vp.gx % 4 == 0)
            hipLaunchKernelGGL(diff_integrate_kernel<true>, grid, dim3(kThreads), 0, s, vp, n, ft,
                               list, h->list_count);
        else
            hipLaunchKernelGGL(diff_integrate_kernel<false>, grid, dim3(kThreads), 0, s, vp, n, ft,
                               list, h->list_count);
        RECON_HIP_TRY(hipGetLastError(), "diff_integrate_kernel launch");
    }
    return RECON_OK;
}

extern "C" int recon_tsdf_surface_loss(recon_tsdf* h, const float* points, const float* weights,
                                       int64_t num_points, float* residuals, double* loss,
                                       int64_t* lambda_ids, float* lambda, recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_surface_loss: handle is NULL");
    if (num_points < 0) return fail(RECON_ERR_SIZE, "recon_tsdf_surface_loss: num_points < 0");
    if (!loss) return fail(RECON_ERR_NULL, "recon_tsdf_surface_loss: loss is NULL");
    if (num_points > 0 && !points) return fail(RECON_ERR_NULL, "recon_tsdf_surface_loss: points NULL");
    if ((lambda_ids == nullptr) != (lambda == nullptr))
        return fail(RECON_ERR_NULL, "recon_tsdf_surface_loss: lambda_ids and lambda go together");
    if (num_points > (int64_t(1) << 40))
        return fail(RECON_ERR_RANGE, "recon_tsdf_surface_loss: too many points");
    int rc = require_gfx950();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const VolumeParams vp = diff_volume_params(h, h->cfg.camera_width > 1 ? h->cfg.camera_width : 2,
                                               h->cfg.camera_height > 1 ? h->cfg.camera_height : 2);
    const int64_t blocks = std::max<int64_t>((num_points + 255) / 256, 1);
    const DiffLayout L = diff_layout(vp, 1, (size_t)blocks);
    if ((rc = ensure_diff_scratch(h, L.total, s))) return rc;
    double* partials = reinterpret_cast<double*>(static_cast<char*>(h->diff_scratch) + L.partial);
    if (num_points > 0) {
        hipLaunchKernelGGL(surface_loss_kernel, dim3((unsigned)blocks), dim3(256), 0, s, vp, points,
                           weights, num_points, residuals, partials, lambda_ids, lambda);
    } else {
        hipLaunchKernelGGL(zero_frames_kernel, dim3(1), dim3(256), 0, s,
                           reinterpret_cast<float*>(partials), 0, 2);
    }
    RECON_HIP_TRY(hipGetLastError(), "surface_loss_kernel launch");
    hipLaunchKernelGGL(fixed_sum_kernel, dim3(1), dim3(256), 0, s, partials, blocks, loss,
                       num_points > 0 ? 1.0 / (double)num_points : 0.0);
    RECON_HIP_TRY(hipGetLastError(), "fixed_sum_kernel launch");
    return RECON_OK;
}

extern "C" int recon_tsdf_diff_adjoint(recon_tsdf* h, const recon_depth_view* depth0,
                                       size_t frame_stride_bytes, int32_t num_frames,
                                       const float* poses, const int64_t* lambda_ids,
                                       const float* lambda, int64_t num_lambda, double* grad_pose,
                                       float* grad_depth, recon_stream_t stream) {
    clear_error();
    int rc = check_frames(h, depth0, frame_stride_bytes, num_frames, poses,
                          "recon_tsdf_diff_adjoint");
    if (rc) return rc;
    if (num_lambda < 0) return fail(RECON_ERR_SIZE, "recon_tsdf_diff_adjoint: num_lambda < 0");
    if (num_lambda > 0 && (!lambda_ids || !lambda))
        return fail(RECON_ERR_NULL, "recon_tsdf_diff_adjoint: lambda is NULL");
    if (!grad_pose && num_frames > 0)
        return fail(RECON_ERR_NULL, "recon_tsdf_diff_adjoint: grad_pose is NULL");
    if (num_frames == 0) return RECON_OK;
    rc = require_gfx950();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const VolumeParams vp = diff_volume_params(h, depth0->cols, depth0->rows);
    const int64_t per_wave = num_frames <= 64 ? 64 / num_frames : 64;  // entries per pass
    const int64_t want_waves = std::max<int64_t>((num_lambda + per_wave - 1) / per_wave, 1);
    int64_t cap = std::max<int64_t>(64, (int64_t(32) << 20) / (48 * (int64_t)num_frames));
    cap = std::min<int64_t>(cap, kAdjointMaxWaves);
    const int64_t blocks = (std::min(want_waves, cap) + 3) / 4;
    const int64_t waves = blocks * 4;
    const DiffLayout L = diff_layout(vp, num_frames, (size_t)waves * 6 * num_frames);
    // repeated voxels are merged first unless ids or counts exceed 32 bits
    const int64_t N = h->info.num_voxels;
    const bool merge = num_lambda > 0 && N < 0xFFFFFFFFll && num_lambda < 0x7FFFFFFFll;
    const MergeLayout M = merge_layout(num_lambda);
    // depth gradients: int64 accumulators (one per depth element of every
    // frame), the block maxima of |dL/dD| and their maximum
    const size_t per_frame = (depth0->step / 4) * (size_t)depth0->rows;
    const int nmax = (int)std::min<int64_t>(std::max<int64_t>((num_lambda + 255) / 256, 1), 1024);
    const size_t gd_base = L.total + (merge ? M.total : 0);
    const size_t gd_blocks = gd_base + align256(8 * per_frame * (size_t)num_frames);
    const size_t gd_maxbits = gd_blocks + align256(4 * (size_t)nmax);
    const size_t need = grad_depth ? gd_maxbits + 256 : gd_base;
    if ((rc = ensure_diff_scratch(h, need, s))) return rc;
    if (merge && (rc = ensure_owner(h, s))) return rc;
    char* base = static_cast<char*>(h->diff_scratch);
    launch_setup(vp, depth0, frame_stride_bytes, num_frames, poses, base, L, s);
    RECON_HIP_TRY(hipGetLastError(), "diff_setup_kernel launch");
    const size_t gstride = num_frames > 1 ? frame_stride_bytes / 4 : 0;
    unsigned long long* gacc = nullptr;
    unsigned* gmax_blocks = nullptr;
    unsigned* gmax = nullptr;
    if (grad_depth) {
        gacc = reinterpret_cast<unsigned long long*>(base + gd_base);
        gmax_blocks = reinterpret_cast<unsigned*>(base + gd_blocks);
        gmax = reinterpret_cast<unsigned*>(base + gd_maxbits);
        const size_t n = per_frame * (size_t)num_frames;
        const unsigned gx = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(zero_u64_kernel, dim3(gx), dim3(256), 0, s, gacc, n);
        RECON_HIP_TRY(hipGetLastError(), "zero_u64_kernel launch");
    }
    double* partial = reinterpret_cast<double*>(base + L.partial);
    const FrameTable ft{reinterpret_cast<const FrameParams*>(base + L.table)};
    const dim3 grid((unsigned)blocks);
    if (merge) {
        // a failure past this point may leave owner entries set: refill later
        h->owner_dirty = true;
        uint32_t* uid = nullptr;
        float* lam_u = nullptr;
        const long long* count = nullptr;
        if ((rc = launch_merge(lambda_ids, lambda, num_lambda, N, h->owner, base + L.total, M, &uid,
                               &lam_u, &count, s)))
            return rc;
        h->owner_dirty = false;
        if (grad_depth) {
            hipLaunchKernelGGL(lambda_max_kernel, dim3(nmax), dim3(256), 0, s, lam_u, num_lambda,
                               count, gmax_blocks);
            hipLaunchKernelGGL(merge_max_kernel, dim3(1), dim3(256), 0, s, gmax_blocks, nmax, gmax);
            RECON_HIP_TRY(hipGetLastError(), "lambda_max_kernel launch");
        }
        if (num_frames <= 64)
            hipLaunchKernelGGL(diff_adjoint_kernel<uint32_t>, grid, dim3(256), 0, s, vp, num_frames,
                               ft, uid, lam_u, num_lambda, count, partial, gacc, per_frame, gmax);
        else
            hipLaunchKernelGGL(diff_adjoint_long_kernel<uint32_t>, grid, dim3(256), 0, s, vp,
                               num_frames, ft, uid, lam_u, num_lambda, count, partial, gacc,
                               per_frame, gmax);
    } else {
        const long long* none = nullptr;
        if (grad_depth) {
            hipLaunchKernelGGL(lambda_max_kernel, dim3(nmax), dim3(256), 0, s, lambda, num_lambda,
                               none, gmax_blocks);
            hipLaunchKernelGGL(merge_max_kernel, dim3(1), dim3(256), 0, s, gmax_blocks, nmax, gmax);
            RECON_HIP_TRY(hipGetLastError(), "lambda_max_kernel launch");
        }
        if (num_frames <= 64)
            hipLaunchKernelGGL(diff_adjoint_kernel<int64_t>, grid, dim3(256), 0, s, vp, num_frames,
                               ft, lambda_ids, lambda, num_lambda, none, partial, gacc, per_frame,
                               gmax);
        else
            hipLaunchKernelGGL(diff_adjoint_long_kernel<int64_t>, grid, dim3(256), 0, s, vp,
                               num_frames, ft, lambda_ids, lambda, num_lambda, none, partial, gacc,
                               per_frame, gmax);
    }
    RECON_HIP_TRY(hipGetLastError(), "diff_adjoint_kernel launch");
    if (grad_depth) {
        const unsigned gx = (unsigned)std::min<size_t>((per_frame + 255) / 256, 1024);
        hipLaunchKernelGGL(depth_grad_finish_kernel, dim3(gx, num_frames), dim3(256), 0, s, gacc,
                           per_frame, grad_depth, gstride, gmax);
        RECON_HIP_TRY(hipGetLastError(), "depth_grad_finish_kernel launch");
    }
    return launch_fixed_sum(partial, waves, grad_pose, 6 * num_frames, 1.0, s);
}
