// diff_batch.hip -- pose refinement over many independent fragments on one
// GPU: the differentiable integrate, the surface loss and the adjoint of
// tsdf_diff.hip over a batch of volumes (sizes, frame counts and image sizes
// may differ), with the poses refined on the device.  The reference has no
// counterpart (it fuses one scene at a time with fixed poses,
// tsdf_surface_reconstruction.cpp:125-138).
//
// A batch owns one arena for every volume's D and W, device tables of
// volumes and frames, the depth bounds of every frame (computed once: depth
// maps do not change), and the scratch of every step, so that a step
// allocates nothing and can be captured.  One step is:
//   forward   frame table from the poses (device), fill of every volume,
//             one cull over the wave-tiles of every volume (64 tiles x 4
//             frame groups per workgroup), one persistent bilinear integrate
//             over the list, the loss of every volume's points (workgroups
//             never straddle volumes, so each volume's loss is its own
//             ordered sum);
//   backward  the (voxel, dL/dD) pairs of all volumes merged (pairs are
//             volume-major, so the merged list is too), work units of up to
//             kUnit merged pairs within one volume, lanes as frames, per-unit
//             partials summed per (frame, component) over the volume's units
//             in order;
//   update    Adam on xi = (rho, omega) per frame with second moments shared
//             per volume (translations, rotations), or a damped Gauss-Newton
//             step per frame from the model-to-frame term's 6x6 normal
//             matrix (accumulated by the forward's frame-term kernel);
//             T <- Exp(step) T in float64, R re-orthonormalised, stored as
//             float32.
// recon_diff_batch_refine and recon_diff_batch_refine_gn capture forward +
// backward + update once and replay the graph per step.
#pragma clang fp contract(off)

#include "diff_device.hpp"
#include "diff_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace recon {

using Shape = DefaultShape;
constexpr int kUnit = 1024;          // merged pairs per adjoint work unit
constexpr int kMaxWaves = 8192;      // bound on the adjoint grid
constexpr int kFill = 1024;          // voxels per fill workgroup

struct DVol {
    VolumeParams vp;
    int64_t voxel_base;      // first voxel of the volume in the arena
    int64_t first_point;     // first point of the volume in the batch
    int64_t npoints;
    const float* points;
    const float* weights;
    int32_t first_frame, nframes;
    int32_t first_loss_block, nloss_blocks;
    const int32_t* fpairs;   // model-to-frame pairs (frame, point), by frame
    int64_t nfpairs;
    double fweight;
    double ftrunc;           // residuals of |r| >= ftrunc count ftrunc^2 (0: off)
    const float* fnormals;   // unit world normals of the points, or null
};

struct DFrame {
    const float* depth;
    const float* bound0;
    const float* bound1;
    int32_t pitch;
    int32_t volume;
};

__device__ __forceinline__ int last_le(const int* starts, int n, int b) {
    int lo = 0, hi = n - 1;  // last i with starts[i] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void dbatch_bounds_kernel(const BoundsJob* __restrict__ jobs) {
    __shared__ int s0[16];
    __shared__ int s1[4];
    const BoundsJob j = jobs[blockIdx.z];
    if ((int)blockIdx.x >= j.b1x || (int)blockIdx.y >= j.b1y) return;  // whole workgroup
    bounds_block(j, blockIdx.x, blockIdx.y, s0, s1);
}

__global__ __launch_bounds__(64) void dbatch_setup_kernel(const DVol* __restrict__ vols,
                                                          const DFrame* __restrict__ fr,
                                                          const float* __restrict__ poses,
                                                          int total_frames,
                                                          FrameParams* __restrict__ table) {
    const int g = blockIdx.x * 64 + threadIdx.x;
    if (g >= total_frames) return;
    const DFrame d = fr[g];
    FrameParams fp;
    for (int i = 0; i < 9; ++i) fp.R[i] = poses[12 * g + i];
    for (int i = 0; i < 3; ++i) fp.t[i] = poses[12 * g + 9 + i];
    fp.depth = d.depth;
    fp.bound0 = d.bound0;
    fp.bound1 = d.bound1;
    fp.pitch = d.pitch;
    frame_planes(vols[d.volume].vp, &fp);
    table[g] = fp;
}

__global__ __launch_bounds__(256) void dbatch_fill_kernel(const DVol* __restrict__ vols,
                                                          const int* __restrict__ block_start,
                                                          int nvol, unsigned* list_count) {
    const int b = blockIdx.x;
    if (b == 0 && threadIdx.x == 0) *list_count = 0;
    const int v = last_le(block_start, nvol, b);
    const VolumeParams& vp = vols[v].vp;
    const long long n = (long long)vp.gx * vp.gy * vp.gz;
    fill4(vp.D, vp.W, (long long)(b - block_start[v]) * kFill + 4 * threadIdx.x, n, vp.trunc);
}

__global__ __launch_bounds__(64 * kCullWaves) void dbatch_cull_kernel(
    const DVol* __restrict__ vols, const FrameParams* __restrict__ table,
    const uint2* __restrict__ wg, ListEntry* __restrict__ list, unsigned* __restrict__ list_count) {
    __shared__ uint64_t s_live[kCullWaves][64];
    const uint2 e = wg[blockIdx.x];  // (volume, first tile)
    const DVol& dv = vols[e.x];
    const uint32_t ntiles = (uint32_t)dv.vp.tiles_x * dv.vp.tiles_y * dv.vp.tiles_z;
    const int i = threadIdx.x & 63, g = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    const uint32_t id = e.y + i;
    s_live[g][i] = id < ntiles ? tile_frames<Shape>(dv.vp, dv.nframes,
                                                    FrameTable{table + dv.first_frame}, id, g, nw)
                               : 0ull;
    __syncthreads();
    if (g != 0) return;
    uint64_t live = 0;
    for (int k = 0; k < nw; ++k) live |= s_live[k][i];
    append_wave(live, id, e.x, list, list_count);
}

__global__ __launch_bounds__(kThreads) void dbatch_integrate_kernel(
    const DVol* __restrict__ vols, const FrameParams* __restrict__ table,
    const ListEntry* __restrict__ list, const unsigned* __restrict__ list_count) {
    __shared__ float rtab[kRcpTable];
    rtab[threadIdx.x] = 1.0f / (float)(threadIdx.x + 1);
    __syncthreads();
    const unsigned n = __builtin_amdgcn_readfirstlane(*list_count);
    const unsigned waves = gridDim.x * (kThreads / 64);
    for (unsigned i = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t tile = __builtin_amdgcn_readfirstlane(list[i].tile);
        const uint32_t v = __builtin_amdgcn_readfirstlane(list[i].volume);
        const uint64_t mask =
            ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(list[i].frames >> 32)) << 32) |
            __builtin_amdgcn_readfirstlane((uint32_t)list[i].frames);
        const DVol& dv = vols[v];
        int x0, y0, z0;
        tile_origin<Shape>(dv.vp, tile, x0, y0, z0);
        const FrameTable ft{table + dv.first_frame};
        if (dv.vp.gx % 4 == 0)
            diff_integrate_tile<Shape, FrameTable, true>(dv.vp, dv.nframes, ft, mask, x0, y0, z0,
                                                         rtab);
        else
            diff_integrate_tile<Shape, FrameTable, false>(dv.vp, dv.nframes, ft, mask, x0, y0, z0,
                                                          rtab);
    }
}

// One workgroup: 256 points of one volume.  Pair k*8 + i of the batch is
// support voxel i of batch point k, as an arena index.
__global__ __launch_bounds__(256) void dbatch_loss_kernel(const DVol* __restrict__ vols,
                                                          const uint2* __restrict__ wg,
                                                          double* __restrict__ partials,
                                                          int64_t* __restrict__ lam_ids,
                                                          float* __restrict__ lam) {
    __shared__ double s_wave[4];
    const uint2 e = wg[blockIdx.x];  // (volume, first local point)
    const DVol& dv = vols[e.x];
    const int64_t k = (int64_t)e.y + threadIdx.x;
    double l = 0.0;
    if (k < dv.npoints) {
        const float p[3] = {dv.points[3 * k], dv.points[3 * k + 1], dv.points[3 * k + 2]};
        const float w = dv.weights ? dv.weights[k] : 1.0f;
        int64_t ids[8];
        float coef[8];
        const bool ok = tri_support(dv.vp, p, ids, coef);
        float r = 0.0f;
        if (ok) {
#pragma unroll
            for (int i = 0; i < 8; ++i) r = r + coef[i] * dv.vp.D[ids[i]];
        }
        l = ok ? (double)w * (double)r * (double)r : 0.0;
        const float g = (2.0f * w * r) / (float)dv.npoints;
        const int64_t q = 8 * (dv.first_point + k);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            lam_ids[q + i] = ok ? dv.voxel_base + ids[i] : -1;
            lam[q + i] = ok ? g * coef[i] : 0.0f;
        }
    }
    l = wave_sum_f64(l);
    if ((threadIdx.x & 63) == 0) s_wave[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = ((s_wave[0] + s_wave[1]) + s_wave[2]) + s_wave[3];
}

// Each volume's loss: its blocks' partials summed in the order of the
// single-volume path's fixed_sum_kernel, times 1/K.
__global__ __launch_bounds__(256) void dbatch_loss_sum_kernel(const DVol* __restrict__ vols,
                                                              const double* __restrict__ partials,
                                                              const double* __restrict__ floss,
                                                              double* __restrict__ loss) {
    __shared__ double s[256];
    const DVol& dv = vols[blockIdx.x];
    const double* src = partials + dv.first_loss_block;
    double acc = 0.0;
    for (int i = threadIdx.x; i < dv.nloss_blocks; i += 256) acc += src[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int m = 128; m >= 1; m >>= 1) {
        if ((int)threadIdx.x < m) s[threadIdx.x] += s[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double l = dv.npoints > 0 ? s[0] * (1.0 / (double)dv.npoints) : 0.0;
        if (floss && dv.nfpairs > 0)
            for (int f = 0; f < dv.nframes; ++f) l += floss[dv.first_frame + f];
        loss[blockIdx.x] = l;
    }
}

// Model-to-frame term: one workgroup per frame g over its pairs
// [fr_range[2g], fr_range[2g+1]) of its volume, in float64 (the frame table's
// float32 R, t widened): the weighted sum of min(r^2, tau^2) / K into
// floss[g] (a truncated residual has no gradient) and its
// gradient w.r.t. the frame's left perturbation into fgrad[6g..6g+5] (the
// adjoint's formula at the point), each summed in a fixed order.  tau is the
// volume's truncation, times ratio^min(1, step / ka) when `anneal` is given
// (Gauss-Newton refinement).  With kHess, the frame's Gauss-Newton system
// instead: fgrad = fweight * 2 sum r J / K and the upper triangle of
// fweight * 2 sum J J^T / K into fhess[21g..21g+20] (row-major), over the
// pairs with |r| < tau, J = (j, p x j): j = -(z / (n . p)) n with n the
// point's normal in the camera frame (point-to-plane; |n . p| < 0.1 |n| |p|
// adds nothing) when the volume has normals, else the bilinear slopes'
// dr/dp (then fgrad is the exact gradient, as without kHess).
struct FrameAnneal {
    const int* step;  // null: fixed truncation
    double ratio, ka;
};

template <bool kHess>
__global__ __launch_bounds__(256) void dbatch_frame_loss_kernel(
    const DVol* __restrict__ vols, const DFrame* __restrict__ fr,
    const FrameParams* __restrict__ table, const int64_t* __restrict__ fr_range,
    const FrameAnneal anneal, double* __restrict__ floss, double* __restrict__ fgrad,
    double* __restrict__ fhess) {
    constexpr int kAcc = kHess ? 28 : 7;
    __shared__ double s_red[kAcc][4];
    const int g = blockIdx.x;
    const DVol& dv = vols[fr[g].volume];
    const VolumeParams& vp = dv.vp;
    const FrameParams& L = table[g];
    const int64_t k0 = fr_range[2 * g], k1 = fr_range[2 * g + 1];
    const double scale = dv.nfpairs > 0 ? dv.fweight / (double)dv.nfpairs : 0.0;
    double tau = dv.ftrunc;
    if (anneal.step) tau *= pow(anneal.ratio, fmin(1.0, (double)*anneal.step / anneal.ka));
    const float* nw = kHess ? dv.fnormals : nullptr;
    double acc[kAcc];
#pragma unroll
    for (int c = 0; c < kAcc; ++c) acc[c] = 0.0;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += 256) {
        const int64_t q = dv.fpairs[2 * k + 1];
        const double c[3] = {(double)dv.points[3 * q], (double)dv.points[3 * q + 1],
                             (double)dv.points[3 * q + 2]};
        double pc[3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
            pc[r] = (double)L.R[3 * r] * c[0] + (double)L.R[3 * r + 1] * c[1] +
                    (double)L.R[3 * r + 2] * c[2] + (double)L.t[r];
        if (!(pc[2] > 0.0)) continue;
        const double u = pc[0] * (double)vp.fx / pc[2] + (double)vp.cx;
        const double v = pc[1] * (double)vp.fy / pc[2] + (double)vp.cy;
        const double fu = floor(u), fv = floor(v);
        if (!(fu >= 0.0 && fu <= vp.cols - 2 && fv >= 0.0 && fv <= vp.rows - 2)) continue;
        const int i0 = (int)fu, j0 = (int)fv;
        const double a = u - fu, b = v - fv;
        const float* t0 = L.depth + (int64_t)j0 * L.pitch + i0;
        const double d00 = t0[0], d10 = t0[1], d01 = t0[L.pitch], d11 = t0[L.pitch + 1];
        const double maxd = vp.maxd;
        if (!(d00 > 0.0 && d00 <= maxd && d10 > 0.0 && d10 <= maxd && d01 > 0.0 && d01 <= maxd &&
              d11 > 0.0 && d11 <= maxd))
            continue;
        const double d = (1 - a) * (1 - b) * d00 + a * (1 - b) * d10 + (1 - a) * b * d01 + a * b * d11;
        const double res = d - pc[2];
        if (tau > 0.0 && !(fabs(res) < tau)) {
            acc[0] += tau * tau;
            continue;
        }
        acc[0] += res * res;
        const double gs = 2.0 * res * scale;
        const double dd_du = (1 - b) * (d10 - d00) + b * (d11 - d01);
        const double dd_dv = (1 - a) * (d01 - d00) + a * (d11 - d10);
        const double iz = 1.0 / pc[2];
        if constexpr (!kHess) {
            const double gx = gs * dd_du * (double)vp.fx * iz;
            const double gy = gs * dd_dv * (double)vp.fy * iz;
            const double gz = -(gx * pc[0] + gy * pc[1]) * iz - gs;
            acc[1] += gx;
            acc[2] += gy;
            acc[3] += gz;
            acc[4] += pc[1] * gz - pc[2] * gy;
            acc[5] += pc[2] * gx - pc[0] * gz;
            acc[6] += pc[0] * gy - pc[1] * gx;
        } else {
            // the Gauss-Newton system of one linearisation: J from the
            // normal (point-to-plane) when the volume has normals, else
            // from the slopes
            double j[3];
            bool use = true;
            if (nw) {
                const double nwx = nw[3 * q], nwy = nw[3 * q + 1], nwz = nw[3 * q + 2];
                double n[3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    n[r] = (double)L.R[3 * r] * nwx + (double)L.R[3 * r + 1] * nwy +
                           (double)L.R[3 * r + 2] * nwz;
                const double np = n[0] * pc[0] + n[1] * pc[1] + n[2] * pc[2];
                const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                const double pn = sqrt(pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
                use = fabs(np) >= 0.1 * nn * pn;
                const double cj = use ? -pc[2] / np : 0.0;
#pragma unroll
                for (int r = 0; r < 3; ++r) j[r] = cj * n[r];
            } else {
                j[0] = dd_du * (double)vp.fx * iz;
                j[1] = dd_dv * (double)vp.fy * iz;
                j[2] = -(j[0] * pc[0] + j[1] * pc[1]) * iz - 1.0;
            }
            if (use) {
                const double J[6] = {j[0], j[1], j[2], pc[1] * j[2] - pc[2] * j[1],
                                     pc[2] * j[0] - pc[0] * j[2], pc[0] * j[1] - pc[1] * j[0]};
#pragma unroll
                for (int r = 0; r < 6; ++r) acc[1 + r] += gs * J[r];
                int e = 7;
#pragma unroll
                for (int r = 0; r < 6; ++r)
#pragma unroll
                    for (int cc = r; cc < 6; ++cc) acc[e++] += J[r] * J[cc];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < kAcc; ++c) {
        const double w = wave_sum_f64(acc[c]);
        if ((threadIdx.x & 63) == 0) s_red[c][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < kAcc) {
        const int c = threadIdx.x;
        const double sum = ((s_red[c][0] + s_red[c][1]) + s_red[c][2]) + s_red[c][3];
        if (c == 0) floss[g] = sum * scale;
        else if (c < 7) fgrad[6 * g + c - 1] = sum;
        else fhess[21 * g + c - 7] = 2.0 * scale * sum;
    }
}

// Per volume, the range of the merged list it owns and its work units.
// One workgroup; thread 0 forms the prefix (volume counts are small).
__global__ __launch_bounds__(256) void dbatch_units_kernel(const DVol* __restrict__ vols, int nvol,
                                                           const uint32_t* __restrict__ uid,
                                                           const long long* __restrict__ count,
                                                           long long* __restrict__ beg,
                                                           long long* __restrict__ unit_start) {
    const long long n = *count;
    for (int v = threadIdx.x; v < nvol; v += 256) {
        // first merged pair at or past the volume's first voxel (pairs are
        // grouped by volume in volume order)
        const long long base = vols[v].voxel_base;
        long long lo = 0, hi = n;
        while (lo < hi) {
            const long long mid = (lo + hi) >> 1;
            if ((long long)uid[mid] < base) lo = mid + 1; else hi = mid;
        }
        beg[v] = lo;
    }
    if (threadIdx.x == 0) beg[nvol] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long u = 0;
        for (int v = 0; v < nvol; ++v) {
            unit_start[v] = u;
            u += (beg[v + 1] - beg[v] + kUnit - 1) / kUnit;
        }
        unit_start[nvol] = u;
    }
}

__global__ __launch_bounds__(256) void dbatch_adjoint_kernel(
    const DVol* __restrict__ vols, int nvol, const FrameParams* __restrict__ table,
    const uint32_t* __restrict__ uid, const float* __restrict__ lam_u,
    const long long* __restrict__ beg, const long long* __restrict__ unit_start,
    double* __restrict__ partial) {
    const unsigned waves = gridDim.x * 4u;
    const unsigned wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const long long nunits = unit_start[nvol];
    for (long long u = wave; u < nunits; u += waves) {
        int lo = 0, hi = nvol - 1;  // last volume whose first unit <= u
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (unit_start[mid] <= u) lo = mid; else hi = mid - 1;
        }
        const int v = __builtin_amdgcn_readfirstlane(lo);
        const DVol& dv = vols[v];
        const LaneLayout ly = lane_layout(dv.nframes);
        const FrameParams L = table[dv.first_frame + (ly.ok ? ly.f : 0)];
        const long long e_begin = beg[v] + (u - unit_start[v]) * kUnit;
        const long long e_end = min(e_begin + kUnit, beg[v + 1]);
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (long long e0 = e_begin; e0 < e_end; e0 += ly.S) {
            const long long e = e0 + ly.seg;
            int64_t vx = -1;
            float l = 0.0f;
            if (ly.ok && e < e_end) {
                vx = (int64_t)uid[e] - dv.voxel_base;
                l = lam_u[e];
            }
            adjoint_lane_pass(dv.vp, ly, L, vx, l, nullptr, 1.0, acc);
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double sum = segment_sum(acc[c], ly);
            if (ly.seg == 0 && ly.ok) partial[((size_t)u * 64 + ly.f) * 6 + c] = sum;
        }
    }
}

// grad[g*6 + c] = sum over the units of frame g's volume, in order.
__global__ __launch_bounds__(256) void dbatch_grad_kernel(const DVol* __restrict__ vols,
                                                          const DFrame* __restrict__ fr,
                                                          int total_frames,
                                                          const long long* __restrict__ unit_start,
                                                          const double* __restrict__ partial,
                                                          const double* __restrict__ fgrad,
                                                          double* __restrict__ grad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= total_frames * 6) return;
    const int g = i / 6, c = i - 6 * g;
    const int v = fr[g].volume;
    const int f = g - vols[v].first_frame;
    double s = 0.0;
    if (unit_start)
        for (long long u = unit_start[v]; u < unit_start[v + 1]; ++u)
            s += partial[((size_t)u * 64 + f) * 6 + c];
    if (fgrad) s += fgrad[i];
    grad[i] = s;
}

// ---- the update ----------------------------------------------------------------------
struct AdamState {
    double lr_t, lr_r, b1, b2, eps, decay;
};

__device__ void se3_exp_f64(const double xi[6], double R[9], double t[3]) {
    const double w0 = xi[3], w1 = xi[4], w2 = xi[5];
    const double th2 = w0 * w0 + w1 * w1 + w2 * w2, th = sqrt(th2);
    double A, B, C;
    if (th < 1e-6) {
        A = 1.0 - th2 / 6.0;
        B = 0.5 - th2 / 24.0;
        C = 1.0 / 6.0 - th2 / 120.0;
    } else {
        A = sin(th) / th;
        B = (1.0 - cos(th)) / th2;
        C = (th - sin(th)) / (th2 * th);
    }
    const double K[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
    double K2[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            K2[3 * r + c] = K[3 * r] * K[c] + K[3 * r + 1] * K[3 + c] + K[3 * r + 2] * K[6 + c];
    double V[9];
    for (int k = 0; k < 9; ++k) {
        const double I = (k % 4 == 0) ? 1.0 : 0.0;
        R[k] = I + A * K[k] + B * K2[k];
        V[k] = I + B * K[k] + C * K2[k];
    }
    for (int r = 0; r < 3; ++r) t[r] = V[3 * r] * xi[0] + V[3 * r + 1] * xi[1] + V[3 * r + 2] * xi[2];
}

// T <- Exp(xi) T in float64 for a float32 [R | t] (12 floats), R's rows
// re-orthonormalised (Gram-Schmidt, third row from the cross product).
__device__ void left_update(float* __restrict__ pose, const double xi[6]) {
    double Re[9], te[3];
    se3_exp_f64(xi, Re, te);
    double R0[9], t0[3], R[9], tt[3];
    for (int k = 0; k < 9; ++k) R0[k] = pose[k];
    for (int k = 0; k < 3; ++k) t0[k] = pose[9 + k];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            R[3 * r + c] = Re[3 * r] * R0[c] + Re[3 * r + 1] * R0[3 + c] + Re[3 * r + 2] * R0[6 + c];
        tt[r] = Re[3 * r] * t0[0] + Re[3 * r + 1] * t0[1] + Re[3 * r + 2] * t0[2] + te[r];
    }
    double* a = R;
    double* b = R + 3;
    double* c = R + 6;
    const double na = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int k = 0; k < 3; ++k) a[k] /= na;
    const double d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    for (int k = 0; k < 3; ++k) b[k] -= d * a[k];
    const double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    for (int k = 0; k < 3; ++k) b[k] /= nb;
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    for (int k = 0; k < 9; ++k) pose[k] = (float)R[k];
    for (int k = 0; k < 3; ++k) pose[9 + k] = (float)tt[k];
}

// One workgroup per volume, one thread per frame.  Adam with a first moment
// per coordinate and a second moment shared by the volume's translations and
// one shared by its rotations (the mean square over the volume's frames): a
// coordinate the loss barely constrains keeps its small step instead of
// being normalised up to the full rate.  step = -lr * mhat / (sqrt(vhat) +
// eps), lr_t for rho and lr_r for omega, both times decay^(t-1);
// T <- Exp(step) T in float64, R re-orthonormalised.
__global__ __launch_bounds__(64) void dbatch_adam_kernel(const DVol* __restrict__ vols,
                                                         float* __restrict__ poses,
                                                         const double* __restrict__ grad,
                                                         double* __restrict__ m,
                                                         double* __restrict__ v2,
                                                         const int* __restrict__ step,
                                                         const AdamState opt) {
    __shared__ double s_sq[2][64];
    const DVol& dv = vols[blockIdx.x];
    const int f = threadIdx.x;
    const bool ok = f < dv.nframes;
    const int g = dv.first_frame + (ok ? f : 0);
    double gk[6];
    double sq_t = 0.0, sq_r = 0.0;
    for (int k = 0; k < 6; ++k) {
        gk[k] = ok ? grad[6 * g + k] : 0.0;
        if (k < 3) sq_t += gk[k] * gk[k]; else sq_r += gk[k] * gk[k];
    }
    s_sq[0][f] = sq_t;
    s_sq[1][f] = sq_r;
    __syncthreads();
    for (int h = 32; h >= 1; h >>= 1) {
        if (f < h) {
            s_sq[0][f] += s_sq[0][f + h];
            s_sq[1][f] += s_sq[1][f + h];
        }
        __syncthreads();
    }
    const double denom = 3.0 * (double)dv.nframes;
    const int t = *step + 1;
    const double c1 = 1.0 - pow(opt.b1, (double)t), c2 = 1.0 - pow(opt.b2, (double)t);
    const double scale = pow(opt.decay, (double)(t - 1));
    // shared second moments live in the slots of the volume's first frame
    double* vs = v2 + 6 * (size_t)dv.first_frame;
    const double vt = opt.b2 * vs[0] + (1.0 - opt.b2) * (s_sq[0][0] / denom);
    const double vr = opt.b2 * vs[1] + (1.0 - opt.b2) * (s_sq[1][0] / denom);
    __syncthreads();
    if (f == 0) {
        vs[0] = vt;
        vs[1] = vr;
    }
    if (!ok) return;
    double xi[6];
    for (int k = 0; k < 6; ++k) {
        const double mk = opt.b1 * m[6 * g + k] + (1.0 - opt.b1) * gk[k];
        m[6 * g + k] = mk;
        const double lr = (k < 3 ? opt.lr_t : opt.lr_r) * scale;
        const double vk = k < 3 ? vt : vr;
        xi[k] = -lr * (mk / c1) / (sqrt(vk / c2) + opt.eps);
    }
    left_update(poses + 12 * (size_t)g, xi);
}

// One thread per frame: a damped Gauss-Newton step.  M = H + mu diag(H) +
// 1e-12 max(diag H) I from the frame's normal matrix (upper triangle, 21
// doubles), factored M = L L^T column by column; a column whose pivot is at
// most 1e-9 max(diag M) (a coordinate the pairs do not fix beyond the ones
// before it) is skipped: its row and column of L are zero and its coordinate
// of x is 0.  x solves the rest; T <- Exp(-alpha x) T.  A frame without
// pairs (H = 0) keeps its pose.
struct GnState {
    double alpha, mu;
};

__global__ __launch_bounds__(64) void dbatch_gn_kernel(int nframes, float* __restrict__ poses,
                                                       const double* __restrict__ grad,
                                                       const double* __restrict__ hess,
                                                       const GnState opt) {
    const int g = blockIdx.x * 64 + threadIdx.x;
    if (g >= nframes) return;
    double M[6][6];
    int e = 0;
    for (int r = 0; r < 6; ++r)
        for (int c = r; c < 6; ++c) {
            M[r][c] = hess[21 * (size_t)g + e++];
            M[c][r] = M[r][c];
        }
    double dmax = 0.0;
    for (int r = 0; r < 6; ++r) dmax = fmax(dmax, M[r][r]);
    if (!(dmax > 0.0)) return;
    for (int r = 0; r < 6; ++r) M[r][r] = M[r][r] * (1.0 + opt.mu) + 1e-12 * dmax;
    double mmax = 0.0;
    for (int r = 0; r < 6; ++r) mmax = fmax(mmax, M[r][r]);
    bool live[6];
    for (int c = 0; c < 6; ++c) {
        double d = M[c][c];
        for (int k = 0; k < c; ++k) d -= M[c][k] * M[c][k];
        live[c] = d > 1e-9 * mmax;
        const double l = live[c] ? sqrt(d) : 0.0;
        M[c][c] = l;
        for (int r = c + 1; r < 6; ++r) {
            double v = M[r][c];
            for (int k = 0; k < c; ++k) v -= M[r][k] * M[c][k];
            M[r][c] = live[c] ? v / l : 0.0;
        }
    }
    double y[6], x[6];
    for (int r = 0; r < 6; ++r) {
        double v = grad[6 * (size_t)g + r];
        for (int k = 0; k < r; ++k) v -= M[r][k] * y[k];
        y[r] = live[r] ? v / M[r][r] : 0.0;
    }
    for (int r = 5; r >= 0; --r) {
        double v = y[r];
        for (int k = r + 1; k < 6; ++k) v -= M[k][r] * x[k];
        x[r] = live[r] ? v / M[r][r] : 0.0;
    }
    double xi[6];
    for (int k = 0; k < 6; ++k) xi[k] = -opt.alpha * x[k];
    left_update(poses + 12 * (size_t)g, xi);
}

// Records the step's losses and advances the step counter.
__global__ __launch_bounds__(256) void dbatch_step_kernel(int* __restrict__ step,
                                                          const double* __restrict__ loss, int nvol,
                                                          double* __restrict__ history) {
    const int t = *step;
    for (int v = threadIdx.x; v < nvol; v += 256)
        if (history) history[(size_t)t * nvol + v] = loss[v];
    __syncthreads();
    if (threadIdx.x == 0) *step = t + 1;
}

__global__ __launch_bounds__(64) void dbatch_reset_step_kernel(int* __restrict__ step) {
    if (threadIdx.x == 0) *step = 0;
}

__global__ __launch_bounds__(256) void dbatch_zero_f64_kernel(double* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        p[i] = 0.0;
}

}  // namespace recon

using namespace recon;

struct recon_diff_batch {
    int n = 0;
    int total_frames = 0;
    int64_t total_points = 0, arena_voxels = 0, npairs = 0;
    std::vector<recon_tsdf_info> info;
    std::vector<DVol> h_vols;
    std::vector<void*> allocs;
    float* arena = nullptr;
    DVol* d_vols = nullptr;
    DFrame* d_frames = nullptr;
    FrameParams* table = nullptr;
    int* d_fill_start = nullptr;
    int fill_blocks = 0;
    uint2* d_cull_wg = nullptr;
    int cull_blocks = 0;
    int cull_nw = 1;  // wavefronts per cull workgroup (recon::cull_waves)
    ListEntry* list = nullptr;
    unsigned* list_count = nullptr;
    int integrate_grid = 0;
    uint2* d_loss_wg = nullptr;
    int loss_blocks = 0;
    double* loss_partials = nullptr;
    int64_t* lam_ids = nullptr;
    float* lam = nullptr;
    double* loss = nullptr;  // internal per-volume loss of refinement steps
    int64_t* fr_range = nullptr;  // model-to-frame pairs of each frame
    double* floss = nullptr;      // per-frame model-to-frame loss and gradient
    double* fgrad = nullptr;
    double* fhess = nullptr;      // per-frame normal matrices (Gauss-Newton refinement)
    bool frame_term = false;
    bool all_frame_terms = false; // every volume has pairs
    // set while a Gauss-Newton step is captured: the annealed truncation and
    // the normal matrices of the frame term
    FrameAnneal anneal{nullptr, 1.0, 1.0};
    bool want_hess = false;
    uint32_t* owner = nullptr;
    char* merge_scratch = nullptr;
    MergeLayout M;
    long long* beg = nullptr;
    long long* unit_start = nullptr;
    int64_t units_bound = 0;
    int adjoint_grid = 0;
    double* partial = nullptr;
    double* grad = nullptr;
    double* adam_m = nullptr;
    double* adam_v = nullptr;
    int* step = nullptr;

    ~recon_diff_batch() {
        for (void* p : allocs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** p, size_t count, const char* what) {
        void* q = nullptr;
        RECON_HIP_TRY(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)), what);
        allocs.push_back(q);
        *p = static_cast<T*>(q);
        return RECON_OK;
    }
    int forward(const float* poses, double* loss_out, hipStream_t s);
    int backward(double* grad_out, hipStream_t s);
    int update(float* poses, const AdamState& opt, double* history, hipStream_t s);
    int update_gn(float* poses, const GnState& opt, double* history, hipStream_t s);
};

int recon_diff_batch::forward(const float* poses, double* loss_out, hipStream_t s) {
    hipLaunchKernelGGL(dbatch_setup_kernel, dim3((total_frames + 63) / 64), dim3(64), 0, s, d_vols,
                       d_frames, poses, total_frames, table);
    hipLaunchKernelGGL(dbatch_fill_kernel, dim3(fill_blocks), dim3(256), 0, s, d_vols, d_fill_start,
                       n, list_count);
    hipLaunchKernelGGL(dbatch_cull_kernel, dim3(cull_blocks), dim3(64 * cull_nw), 0, s, d_vols,
                       table, d_cull_wg, list, list_count);
    hipLaunchKernelGGL(dbatch_integrate_kernel, dim3(integrate_grid), dim3(kThreads), 0, s, d_vols,
                       table, list, list_count);
    RECON_HIP_TRY(hipGetLastError(), "diff batch forward launch");
    if (loss_blocks > 0)
        hipLaunchKernelGGL(dbatch_loss_kernel, dim3(loss_blocks), dim3(256), 0, s, d_vols,
                           d_loss_wg, loss_partials, lam_ids, lam);
    if (frame_term && want_hess)
        hipLaunchKernelGGL(dbatch_frame_loss_kernel<true>, dim3(total_frames), dim3(256), 0, s,
                           d_vols, d_frames, table, fr_range, anneal, floss, fgrad, fhess);
    else if (frame_term)
        hipLaunchKernelGGL(dbatch_frame_loss_kernel<false>, dim3(total_frames), dim3(256), 0, s,
                           d_vols, d_frames, table, fr_range, anneal, floss, fgrad, nullptr);
    hipLaunchKernelGGL(dbatch_loss_sum_kernel, dim3(n), dim3(256), 0, s, d_vols, loss_partials,
                       frame_term ? floss : nullptr, loss_out);
    RECON_HIP_TRY(hipGetLastError(), "diff batch loss launch");
    return RECON_OK;
}

int recon_diff_batch::backward(double* grad_out, hipStream_t s) {
    if (npairs == 0) {
        hipLaunchKernelGGL(dbatch_grad_kernel, dim3((total_frames * 6 + 255) / 256), dim3(256), 0,
                           s, d_vols, d_frames, total_frames, nullptr, nullptr,
                           frame_term ? fgrad : nullptr, grad_out);
        RECON_HIP_TRY(hipGetLastError(), "diff batch gradient launch");
        return RECON_OK;
    }
    uint32_t* uid = nullptr;
    float* lam_u = nullptr;
    const long long* count = nullptr;
    int rc = launch_merge(lam_ids, lam, npairs, arena_voxels, owner, merge_scratch, M, &uid, &lam_u,
                          &count, s);
    if (rc) return rc;
    hipLaunchKernelGGL(dbatch_units_kernel, dim3(1), dim3(256), 0, s, d_vols, n, uid, count, beg,
                       unit_start);
    hipLaunchKernelGGL(dbatch_adjoint_kernel, dim3(adjoint_grid), dim3(256), 0, s, d_vols, n, table,
                       uid, lam_u, beg, unit_start, partial);
    hipLaunchKernelGGL(dbatch_grad_kernel, dim3((total_frames * 6 + 255) / 256), dim3(256), 0, s,
                       d_vols, d_frames, total_frames, unit_start, partial,
                       frame_term ? fgrad : nullptr, grad_out);
    RECON_HIP_TRY(hipGetLastError(), "diff batch backward launch");
    return RECON_OK;
}

int recon_diff_batch::update(float* poses, const AdamState& opt, double* history, hipStream_t s) {
    hipLaunchKernelGGL(dbatch_adam_kernel, dim3(n), dim3(64), 0, s, d_vols, poses, grad, adam_m,
                       adam_v, step, opt);
    hipLaunchKernelGGL(dbatch_step_kernel, dim3(1), dim3(256), 0, s, step, loss, n, history);
    RECON_HIP_TRY(hipGetLastError(), "diff batch update launch");
    return RECON_OK;
}

int recon_diff_batch::update_gn(float* poses, const GnState& opt, double* history, hipStream_t s) {
    hipLaunchKernelGGL(dbatch_gn_kernel, dim3((total_frames + 63) / 64), dim3(64), 0, s,
                       total_frames, poses, grad, fhess, opt);
    hipLaunchKernelGGL(dbatch_step_kernel, dim3(1), dim3(256), 0, s, step, loss, n, history);
    RECON_HIP_TRY(hipGetLastError(), "diff batch Gauss-Newton update launch");
    return RECON_OK;
}

static int check_dvolume(const recon_diff_volume& v, int i, recon_tsdf_info* info) {
    int rc = describe(&v.config, info);
    if (rc) return rc;
    if (info->weight_type != RECON_WEIGHT_BY_DEPTH)
        return fail(RECON_ERR_CONFIG, "recon_diff_batch_create: volume %d: the differentiable "
                    "integrate weighs every observation 1 (WEIGHT_BY_DEPTH)", i);
    if (v.num_frames < 1 || v.num_frames > 64)
        return fail(RECON_ERR_SIZE, "recon_diff_batch_create: volume %d has %d frames (1..64)", i,
                    v.num_frames);
    if (v.rows < 2 || v.cols < 2)
        return fail(RECON_ERR_SIZE, "recon_diff_batch_create: volume %d image %dx%d (>= 2x2)", i,
                    v.cols, v.rows);
    if (!v.depth) return fail(RECON_ERR_NULL, "recon_diff_batch_create: volume %d depth is NULL", i);
    if (v.depth_step < 4 * (size_t)v.cols || v.depth_step % 4)
        return fail(RECON_ERR_SIZE, "recon_diff_batch_create: volume %d depth step %zu", i,
                    v.depth_step);
    if (v.num_frames > 1 && (v.frame_stride % 4 || v.frame_stride < v.depth_step * (size_t)v.rows))
        return fail(RECON_ERR_SIZE, "recon_diff_batch_create: volume %d frame stride %zu", i,
                    v.frame_stride);
    if ((v.depth_step / 4) * (size_t)v.rows >= (size_t(1) << 31))
        return fail(RECON_ERR_RANGE, "recon_diff_batch_create: volume %d image too large", i);
    if (v.num_points < 0 || (v.num_points > 0 && !v.points))
        return fail(RECON_ERR_NULL, "recon_diff_batch_create: volume %d points", i);
    if (v.num_points >= (int64_t(1) << 32))
        return fail(RECON_ERR_RANGE, "recon_diff_batch_create: volume %d has too many points", i);
    if (v.num_frame_pairs < 0 || (v.num_frame_pairs > 0 && !v.frame_pairs))
        return fail(RECON_ERR_NULL, "recon_diff_batch_create: volume %d frame pairs", i);
    if (v.num_frame_pairs > 0 && !(v.frame_weight >= 0.0 && v.frame_weight < 1e30))
        return fail(RECON_ERR_CONFIG, "recon_diff_batch_create: volume %d frame weight", i);
    return RECON_OK;
}

static int plan(const recon_diff_volume* vols, int n, recon_diff_batch* b, hipStream_t s) {
    int rc;
    b->n = n;
    b->info.resize(n);
    int64_t arena = 0, frames = 0, points = 0;
    for (int i = 0; i < n; ++i) {
        if ((rc = check_dvolume(vols[i], i, &b->info[i]))) return rc;
        arena += b->info[i].num_voxels;
        frames += vols[i].num_frames;
        points += vols[i].num_points;
    }
    if (arena >= 0xFFFFFFFFll)
        return fail(RECON_ERR_RANGE, "recon_diff_batch_create: %lld voxels exceed 2^32 - 1",
                    (long long)arena);
    if (8 * points >= 0x7FFFFFFFll)
        return fail(RECON_ERR_RANGE, "recon_diff_batch_create: %lld points exceed 2^28",
                    (long long)points);
    b->arena_voxels = arena;
    b->all_frame_terms = n > 0;
    b->total_frames = (int)frames;
    b->total_points = points;
    b->npairs = 8 * points;
    if ((rc = b->alloc(&b->arena, 2 * (size_t)arena, "diff batch arena"))) return rc;

    // volume and frame tables, depth bounds jobs, fill / cull / loss work tables
    std::vector<DFrame> h_frames;
    std::vector<BoundsJob> jobs;
    std::vector<int> fill_start;
    std::vector<uint2> cull_wg, loss_wg;
    std::vector<int64_t> fr_ranges;
    size_t bounds_floats = 0;
    for (int i = 0; i < n; ++i) {
        const recon_diff_volume& v = vols[i];
        const int cols = v.cols, rows = v.rows;
        bounds_floats += (size_t)v.num_frames *
                         ((size_t)((cols + kBound0 - 1) / kBound0) * ((rows + kBound0 - 1) / kBound0) +
                          (size_t)((cols + kBound1 - 1) / kBound1) * ((rows + kBound1 - 1) / kBound1));
    }
    float* bounds = nullptr;
    if ((rc = b->alloc(&bounds, bounds_floats, "diff batch bounds"))) return rc;
    int64_t vbase = 0, pbase = 0;
    size_t boff = 0;
    int fbase = 0, fill_blocks = 0, loss_blocks = 0, max_b1x = 1, max_b1y = 1;
    for (int i = 0; i < n; ++i) {
        const recon_diff_volume& v = vols[i];
        const recon_tsdf_info& inf = b->info[i];
        DVol dv;
        std::memset(&dv, 0, sizeof(dv));
        VolumeParams& vp = dv.vp;
        vp.D = b->arena + vbase;
        vp.W = b->arena + arena + vbase;
        for (int k = 0; k < 3; ++k) {
            vp.vs[k] = inf.voxel_size[k];
            vp.off[k] = v.config.global_offset[k];
        }
        vp.gx = v.config.voxel_grid_size[0];
        vp.gy = v.config.voxel_grid_size[1];
        vp.gz = v.config.voxel_grid_size[2];
        vp.tiles_x = (vp.gx + Shape::X - 1) / Shape::X;
        vp.tiles_y = (vp.gy + Shape::Y - 1) / Shape::Y;
        vp.tiles_z = (vp.gz + Shape::Z - 1) / Shape::Z;
        vp.fx = v.config.focal_x;
        vp.fy = v.config.focal_y;
        vp.cx = v.config.cx;
        vp.cy = v.config.cy;
        vp.trunc = inf.truncation;
        vp.maxw = v.config.max_weight;
        vp.maxd = inf.max_depth;
        vp.cols = v.cols;
        vp.rows = v.rows;
        vp.fcols = (float)v.cols;
        vp.frows = (float)v.rows;
        vp.b0x = (v.cols + kBound0 - 1) / kBound0;
        vp.b0y = (v.rows + kBound0 - 1) / kBound0;
        vp.b1x = (v.cols + kBound1 - 1) / kBound1;
        vp.b1y = (v.rows + kBound1 - 1) / kBound1;
        max_b1x = std::max(max_b1x, vp.b1x);
        max_b1y = std::max(max_b1y, vp.b1y);
        dv.voxel_base = vbase;
        dv.first_point = pbase;
        dv.npoints = v.num_points;
        dv.points = v.points;
        dv.weights = v.weights;
        dv.first_frame = fbase;
        dv.nframes = v.num_frames;
        dv.first_loss_block = loss_blocks;
        dv.nloss_blocks = (int)((v.num_points + 255) / 256);
        dv.fpairs = v.num_frame_pairs > 0 ? v.frame_pairs : nullptr;
        dv.nfpairs = v.num_frame_pairs;
        dv.fweight = v.frame_weight;
        dv.ftrunc = v.frame_trunc > 0.0 ? v.frame_trunc : 0.0;
        dv.fnormals = v.num_frame_pairs > 0 ? v.normals : nullptr;
        if (v.num_frame_pairs == 0) b->all_frame_terms = false;
        // the model-to-frame pairs of each frame: checked on the host (frames
        // grouped in increasing order, points in range), ranges per frame
        std::vector<int64_t> rng(2 * (size_t)v.num_frames, 0);
        if (v.num_frame_pairs > 0) {
            std::vector<int32_t> hp(2 * (size_t)v.num_frame_pairs);
            RECON_HIP_TRY(hipMemcpy(hp.data(), v.frame_pairs, hp.size() * sizeof(int32_t),
                                    hipMemcpyDeviceToHost), "frame pairs download");
            int32_t prev = 0;
            for (int64_t k = 0; k < v.num_frame_pairs; ++k) {
                const int32_t f = hp[2 * k], q = hp[2 * k + 1];
                if (f < prev || f >= v.num_frames || q < 0 || q >= v.num_points)
                    return fail(RECON_ERR_SIZE, "recon_diff_batch_create: volume %d frame pair %lld "
                                "(%d, %d) is out of range or out of frame order", i, (long long)k, f, q);
                if (k == 0 || f != prev) rng[2 * f] = k;
                rng[2 * f + 1] = k + 1;
                prev = f;
            }
            b->frame_term = true;
        }
        fr_ranges.insert(fr_ranges.end(), rng.begin(), rng.end());
        for (int f = 0; f < v.num_frames; ++f) {
            DFrame fr;
            fr.depth = reinterpret_cast<const float*>(reinterpret_cast<const char*>(v.depth) +
                                                      v.frame_stride * (size_t)f);
            fr.bound0 = bounds + boff;
            fr.bound1 = fr.bound0 + (size_t)vp.b0x * vp.b0y;
            fr.pitch = (int32_t)(v.depth_step / 4);
            fr.volume = i;
            h_frames.push_back(fr);
            BoundsJob j;
            j.depth = fr.depth;
            j.bound0 = bounds + boff;
            j.bound1 = j.bound0 + (size_t)vp.b0x * vp.b0y;
            j.pitch = fr.pitch;
            j.cols = v.cols;
            j.rows = v.rows;
            j.b0x = vp.b0x;
            j.b0y = vp.b0y;
            j.b1x = vp.b1x;
            j.b1y = vp.b1y;
            j.maxd = vp.maxd;
            jobs.push_back(j);
            boff += (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
        }
        fill_start.push_back(fill_blocks);
        fill_blocks += (int)((inf.num_voxels + kFill - 1) / kFill);
        const uint32_t ntiles = (uint32_t)vp.tiles_x * vp.tiles_y * vp.tiles_z;
        for (uint32_t t = 0; t < ntiles; t += 64) cull_wg.push_back(make_uint2(i, t));
        for (int64_t p = 0; p < v.num_points; p += 256) loss_wg.push_back(make_uint2(i, (uint32_t)p));
        loss_blocks += dv.nloss_blocks;
        b->h_vols.push_back(dv);
        vbase += inf.num_voxels;
        pbase += v.num_points;
        fbase += v.num_frames;
    }
    b->fill_blocks = fill_blocks;
    b->cull_blocks = (int)cull_wg.size();
    b->loss_blocks = loss_blocks;
    if ((rc = b->alloc(&b->d_vols, n, "diff batch volumes"))) return rc;
    if ((rc = b->alloc(&b->d_frames, h_frames.size(), "diff batch frames"))) return rc;
    if ((rc = b->alloc(&b->table, h_frames.size(), "diff batch frame table"))) return rc;
    BoundsJob* d_jobs = nullptr;
    if ((rc = b->alloc(&d_jobs, jobs.size(), "diff batch bounds jobs"))) return rc;
    if ((rc = b->alloc(&b->d_fill_start, n, "diff batch fill table"))) return rc;
    if ((rc = b->alloc(&b->d_cull_wg, cull_wg.size(), "diff batch cull table"))) return rc;
    if ((rc = b->alloc(&b->d_loss_wg, loss_wg.size(), "diff batch loss table"))) return rc;
    size_t tiles_total = (size_t)cull_wg.size() * 64;
    if ((rc = b->alloc(&b->list, tiles_total, "diff batch cull list"))) return rc;
    if ((rc = b->alloc(&b->list_count, 1, "diff batch list count"))) return rc;
    if ((rc = b->alloc(&b->loss_partials, std::max(loss_blocks, 1), "diff batch loss partials")))
        return rc;
    if ((rc = b->alloc(&b->lam_ids, b->npairs, "diff batch pairs"))) return rc;
    if ((rc = b->alloc(&b->lam, b->npairs, "diff batch pairs"))) return rc;
    if ((rc = b->alloc(&b->loss, n, "diff batch loss"))) return rc;
    if ((rc = b->alloc(&b->fr_range, fr_ranges.size(), "diff batch frame pairs"))) return rc;
    if ((rc = b->alloc(&b->floss, (size_t)frames, "diff batch frame losses"))) return rc;
    if ((rc = b->alloc(&b->fgrad, (size_t)frames * 6, "diff batch frame gradients"))) return rc;
    if ((rc = b->alloc(&b->fhess, (size_t)frames * 21, "diff batch frame normal matrices")))
        return rc;
    if ((rc = b->alloc(&b->owner, arena, "diff batch owner table"))) return rc;
    b->M = merge_layout(b->npairs);
    if ((rc = b->alloc(&b->merge_scratch, b->M.total, "diff batch merge scratch"))) return rc;
    if ((rc = b->alloc(&b->beg, n + 1, "diff batch ranges"))) return rc;
    if ((rc = b->alloc(&b->unit_start, n + 1, "diff batch units"))) return rc;
    b->units_bound = (b->npairs + kUnit - 1) / kUnit + n;
    b->adjoint_grid = (int)((std::min<int64_t>(b->units_bound, kMaxWaves) + 3) / 4);
    if ((rc = b->alloc(&b->partial, (size_t)b->units_bound * 64 * 6, "diff batch partials")))
        return rc;
    if ((rc = b->alloc(&b->grad, (size_t)frames * 6, "diff batch gradients"))) return rc;
    if ((rc = b->alloc(&b->adam_m, (size_t)frames * 6, "diff batch adam"))) return rc;
    if ((rc = b->alloc(&b->adam_v, (size_t)frames * 6, "diff batch adam"))) return rc;
    if ((rc = b->alloc(&b->step, 1, "diff batch step"))) return rc;
    RECON_HIP_TRY(hipMemcpyAsync(b->d_vols, b->h_vols.data(), n * sizeof(DVol),
                                 hipMemcpyHostToDevice, s), "diff batch upload");
    RECON_HIP_TRY(hipMemcpyAsync(b->d_frames, h_frames.data(), h_frames.size() * sizeof(DFrame),
                                 hipMemcpyHostToDevice, s), "diff batch upload");
    RECON_HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(BoundsJob),
                                 hipMemcpyHostToDevice, s), "diff batch upload");
    RECON_HIP_TRY(hipMemcpyAsync(b->d_fill_start, fill_start.data(), n * sizeof(int),
                                 hipMemcpyHostToDevice, s), "diff batch upload");
    RECON_HIP_TRY(hipMemcpyAsync(b->fr_range, fr_ranges.data(), fr_ranges.size() * sizeof(int64_t),
                                 hipMemcpyHostToDevice, s), "diff batch upload");
    RECON_HIP_TRY(hipMemcpyAsync(b->d_cull_wg, cull_wg.data(), cull_wg.size() * sizeof(uint2),
                                 hipMemcpyHostToDevice, s), "diff batch upload");
    if (!loss_wg.empty())
        RECON_HIP_TRY(hipMemcpyAsync(b->d_loss_wg, loss_wg.data(), loss_wg.size() * sizeof(uint2),
                                     hipMemcpyHostToDevice, s), "diff batch upload");
    hipLaunchKernelGGL(dbatch_bounds_kernel, dim3(max_b1x, max_b1y, (unsigned)jobs.size()), dim3(256),
                       0, s, d_jobs);
    RECON_HIP_TRY(hipGetLastError(), "dbatch_bounds_kernel launch");
    if ((rc = launch_fill_u32(b->owner, arena, 0xFFFFFFFFu, s))) return rc;
    int dev = 0, cus = 256, occ = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, dbatch_integrate_kernel, kThreads, 0) !=
            hipSuccess || occ < 1)
        occ = 2;
    b->integrate_grid = cus * occ;
    int most_frames = 1;
    for (const DVol& dv : b->h_vols) most_frames = std::max(most_frames, dv.nframes);
    b->cull_nw = cull_waves(b->cull_blocks, most_frames, cus);
    // the host tables must outlive the copies
    RECON_HIP_TRY(hipStreamSynchronize(s), "diff batch create synchronize");
    return RECON_OK;
}

extern "C" int recon_diff_batch_create(const recon_diff_volume* volumes, int32_t num_volumes,
                                       recon_diff_batch** out) {
    clear_error();
    if (!out) return fail(RECON_ERR_NULL, "recon_diff_batch_create: out is NULL");
    *out = nullptr;
    if (!volumes || num_volumes < 1)
        return fail(RECON_ERR_SIZE, "recon_diff_batch_create: need at least one volume");
    int rc = require_gfx950();
    if (rc) return rc;
    recon_diff_batch* b = new recon_diff_batch();
    rc = plan(volumes, num_volumes, b, nullptr);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return RECON_OK;
}

extern "C" void recon_diff_batch_destroy(recon_diff_batch* b) { delete b; }

extern "C" int recon_diff_batch_volume_data(recon_diff_batch* b, int32_t i, float** distances,
                                            float** weights, recon_tsdf_info* info) {
    clear_error();
    if (!b) return fail(RECON_ERR_NULL, "recon_diff_batch_volume_data: batch is NULL");
    if (i < 0 || i >= b->n) return fail(RECON_ERR_RANGE, "recon_diff_batch_volume_data: index %d", i);
    if (distances) *distances = b->h_vols[i].vp.D;
    if (weights) *weights = b->h_vols[i].vp.W;
    if (info) *info = b->info[i];
    return RECON_OK;
}

extern "C" int recon_diff_batch_download(recon_diff_batch* b, int32_t i, float* host_d,
                                         float* host_w, recon_stream_t stream) {
    clear_error();
    if (!b) return fail(RECON_ERR_NULL, "recon_diff_batch_download: batch is NULL");
    if (i < 0 || i >= b->n) return fail(RECON_ERR_RANGE, "recon_diff_batch_download: index %d", i);
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = (size_t)b->info[i].num_voxels * sizeof(float);
    if (host_d)
        RECON_HIP_TRY(hipMemcpyAsync(host_d, b->h_vols[i].vp.D, bytes, hipMemcpyDeviceToHost, s),
                      "recon_diff_batch_download");
    if (host_w)
        RECON_HIP_TRY(hipMemcpyAsync(host_w, b->h_vols[i].vp.W, bytes, hipMemcpyDeviceToHost, s),
                      "recon_diff_batch_download");
    RECON_HIP_TRY(hipStreamSynchronize(s), "recon_diff_batch_download");
    return RECON_OK;
}

extern "C" int recon_diff_batch_forward(recon_diff_batch* b, const float* poses_rt, double* loss,
                                        recon_stream_t stream) {
    clear_error();
    if (!b || !poses_rt || !loss) return fail(RECON_ERR_NULL, "recon_diff_batch_forward: NULL");
    return b->forward(poses_rt, loss, (hipStream_t)stream);
}

extern "C" int recon_diff_batch_backward(recon_diff_batch* b, double* grad_pose,
                                         recon_stream_t stream) {
    clear_error();
    if (!b || !grad_pose) return fail(RECON_ERR_NULL, "recon_diff_batch_backward: NULL");
    return b->backward(grad_pose, (hipStream_t)stream);
}

// One step (forward, backward, `update`) captured once and replayed `steps`
// times on s.
template <class Update>
static int replay_steps(recon_diff_batch* b, float* poses_rt, int steps, hipStream_t s,
                        Update update) {
    hipStream_t cs = nullptr;
    RECON_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream");
    int rc = RECON_OK;
    const hipError_t e0 = hipStreamSynchronize(s);
    if (e0 != hipSuccess) rc = hip_fail(e0, "refine synchronize");
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (!rc) {
        const hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            rc = b->forward(poses_rt, b->loss, cs);
            if (!rc) rc = b->backward(b->grad, cs);
            if (!rc) rc = update(cs);
            const hipError_t e2 = hipStreamEndCapture(cs, &graph);
            if (!rc && e2 != hipSuccess) rc = hip_fail(e2, "hipStreamEndCapture");
            if (!rc) {
                const hipError_t e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (e3 != hipSuccess) rc = hip_fail(e3, "hipGraphInstantiate");
            }
        } else {
            rc = hip_fail(e, "hipStreamBeginCapture");
        }
    }
    for (int k = 0; !rc && k < steps; ++k) {
        const hipError_t e4 = hipGraphLaunch(exec, s);
        if (e4 != hipSuccess) rc = hip_fail(e4, "hipGraphLaunch");
    }
    // the graph may still run: wait before releasing it
    const hipError_t e5 = hipStreamSynchronize(s);
    if (!rc && e5 != hipSuccess) rc = hip_fail(e5, "refine synchronize");
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cs);
    return rc;
}

extern "C" int recon_diff_batch_refine(recon_diff_batch* b, float* poses_rt, int32_t steps,
                                       const recon_adam_params* opt, double* loss_history,
                                       recon_stream_t stream) {
    clear_error();
    if (!b || !poses_rt || !opt) return fail(RECON_ERR_NULL, "recon_diff_batch_refine: NULL");
    if (steps < 0) return fail(RECON_ERR_SIZE, "recon_diff_batch_refine: steps < 0");
    const AdamState st{opt->lr_translation, opt->lr_rotation, opt->beta1, opt->beta2, opt->eps,
                       opt->decay};
    if (!(st.lr_t >= 0 && st.lr_r >= 0 && st.b1 >= 0 && st.b1 < 1 && st.b2 >= 0 && st.b2 < 1 &&
          st.eps > 0 && st.decay > 0))
        return fail(RECON_ERR_CONFIG, "recon_diff_batch_refine: bad optimiser parameters");
    hipStream_t s = (hipStream_t)stream;
    // fresh optimiser state
    const int64_t nm = (int64_t)b->total_frames * 6;
    hipLaunchKernelGGL(dbatch_zero_f64_kernel, dim3(64), dim3(256), 0, s, b->adam_m, nm);
    hipLaunchKernelGGL(dbatch_zero_f64_kernel, dim3(64), dim3(256), 0, s, b->adam_v, nm);
    hipLaunchKernelGGL(dbatch_reset_step_kernel, dim3(1), dim3(64), 0, s, b->step);
    RECON_HIP_TRY(hipGetLastError(), "diff batch refine reset");
    if (steps == 0) return RECON_OK;
    return replay_steps(b, poses_rt, steps, s, [&](hipStream_t cs) {
        return b->update(poses_rt, st, loss_history, cs);
    });
}

extern "C" int recon_diff_batch_refine_gn(recon_diff_batch* b, float* poses_rt, int32_t steps,
                                          const recon_gn_params* opt, double* loss_history,
                                          recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_diff_batch_refine_gn";
    if (!b || !poses_rt || !opt) return fail(RECON_ERR_NULL, "%s: NULL", fn);
    if (steps < 0) return fail(RECON_ERR_SIZE, "%s: steps < 0", fn);
    if (!(opt->step > 0.0 && opt->step <= 1.0 && opt->damping >= 0.0 && opt->damping < 1e30 &&
          opt->trunc_ratio > 0.0 && opt->trunc_ratio <= 1.0))
        return fail(RECON_ERR_CONFIG, "%s: step must be in (0, 1], damping >= 0 and trunc_ratio "
                    "in (0, 1]", fn);
    if (!b->all_frame_terms)
        return fail(RECON_ERR_CONFIG, "%s: every volume needs model-to-frame pairs", fn);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(dbatch_reset_step_kernel, dim3(1), dim3(64), 0, s, b->step);
    RECON_HIP_TRY(hipGetLastError(), "diff batch refine reset");
    if (steps == 0) return RECON_OK;
    const GnState st{opt->step, opt->damping};
    b->anneal = FrameAnneal{b->step, opt->trunc_ratio, (double)std::max(1, steps / 2 - 1)};
    b->want_hess = true;
    const int rc = replay_steps(b, poses_rt, steps, s, [&](hipStream_t cs) {
        return b->update_gn(poses_rt, st, loss_history, cs);
    });
    b->anneal = FrameAnneal{nullptr, 1.0, 1.0};
    b->want_hess = false;
    return rc;
}
