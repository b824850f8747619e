// tsdf_volume.hip -- TSDF volume lifecycle and the projective integrate
// kernels for gfx950.
//
// Replaces TSDFVolume (cu-Reconstruction modules/src/cuda/tsdf_volume.cu:
// 144-290).  The reference runs one thread per (y,z) voxel row looping over x
// (tsdf_volume.cu:151-163: neighbouring lanes are a row apart, so no access
// coalesces) and re-reads a 28-byte voxel-centre table per voxel.  Here one
// integrate call is three launches on the caller's stream:
//   1. depth_bounds_kernel  per 8x8 and 32x32 pixel block, the largest valid
//                           depth of each frame (and the list counter <- 0);
//   2. cull_kernel          one thread per wave-tile (32 x 8 x 2 voxels):
//                           the tile survives a frame if its voxel-centre box
//                           meets the frame's frustum and the depth bound of
//                           its image footprint is not closer than the box
//                           minus the truncation; survivors are appended to a
//                           compact list with their frame mask;
//   3. integrate_list_kernel  a grid sized to the resident wave slots walks
//                           the list, one wave-tile per wavefront: lane l owns
//                           voxels x..x+3 of row y for two z, so a wave moves
//                           aligned 128-byte row segments of D and W with
//                           dwordx4 accesses, and only lanes with a voxel that
//                           passes read (and write) D and W.
// Voxel centres are computed from the index (no table traffic).  Up to
// kMaxFusedFrames frames can go through one sweep: each voxel sees the frames
// in order, so the result equals one sweep per frame.  The culling is
// conservative (margins far above float32 rounding), so every voxel the
// reference updates is updated here with the same float32 operations.
#pragma clang fp contract(off)

#include "recon_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace recon {

// Wave-tile: 32 x 8 x 2 voxels.  Lane l owns voxels x..x+3 (x = x0 + 4*(l % 8))
// of row y = y0 + l / 8 for z = z0, z0+1: every row segment of a wave is one
// aligned 128-byte line of D and of W when gx is a multiple of 32.
constexpr int kRowsPerLane = 2;
constexpr int kLanesX = 8;
constexpr int kTileX = 4 * kLanesX, kTileY = 64 / kLanesX, kTileZ = kRowsPerLane;
constexpr int kThreads = 256;
constexpr int kCullTilesPerThread = 1;

template <int MAXF>
struct FrameBlock {
    FrameParams f[MAXF];
};

struct ListEntry {
    uint32_t tile;    // wave-tile index x + tiles_x * (y + tiles_y * z)
    uint32_t frames;  // bit f: frame f may update the tile
};

// Max over a voxel-centre box of n.c + k.
__device__ __forceinline__ float plane_max(const float* p, const float lo[3], const float hi[3]) {
    float g = p[3];
    g = g + fmaxf(p[0] * lo[0], p[0] * hi[0]);
    g = g + fmaxf(p[1] * lo[1], p[1] * hi[1]);
    g = g + fmaxf(p[2] * lo[2], p[2] * hi[2]);
    return g;
}

__device__ __forceinline__ void tile_origin(const VolumeParams& vp, uint32_t id, int& x0, int& y0,
                                            int& z0) {
    const uint32_t tx = id % (uint32_t)vp.tiles_x;
    const uint32_t tyz = id / (uint32_t)vp.tiles_x;
    x0 = (int)tx * kTileX;
    y0 = (int)(tyz % (uint32_t)vp.tiles_y) * kTileY;
    z0 = (int)(tyz / (uint32_t)vp.tiles_y) * kTileZ;
}

// True if some block of [bx0,bx1] x [by0,by1] of a bound level holds a depth
// >= need.  Four independent loads per step keep the reads in flight together.
__device__ __forceinline__ bool bound_rect_reaches(const float* __restrict__ lvl, int stride,
                                                   int bx0, int bx1, int by0, int by1,
                                                   float need) {
    for (int by = by0; by <= by1; ++by) {
        const float* row = lvl + by * stride;
        for (int bx = bx0; bx <= bx1; bx += 4) {
            // indices past bx1 repeat bx1, which leaves the maximum unchanged
            const float a = row[bx];
            const float b = row[min(bx + 1, bx1)];
            const float c = row[min(bx + 2, bx1)];
            const float d = row[min(bx + 3, bx1)];
            if (fmaxf(fmaxf(a, b), fmaxf(c, d)) >= need) return true;
        }
    }
    return false;
}

// Depth-bound test of a voxel-centre box for one frame (one thread).
// Returns false only when no voxel of the box can be updated by the frame:
// the box's eight corners bound the camera z of its voxels and, when all lie
// in front of the camera, the hull of their projections bounds the voxels'
// pixel coordinates.  Two pixels of slack cover float32 rounding, the
// hardware reciprocal used here (<= 1 ulp) and the truncation to an index.
__device__ __forceinline__ bool depth_reaches(const VolumeParams& vp, const FrameParams& F,
                                              const float lo[3], const float hi[3]) {
    float zmin = 3.0e38f, umin = 3.0e38f, umax = -3.0e38f, vmin = 3.0e38f, vmax = -3.0e38f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float cx = (i & 1) ? hi[0] : lo[0];
        const float cy = (i & 2) ? hi[1] : lo[1];
        const float cz = (i & 4) ? hi[2] : lo[2];
        const float px = ((F.R[0] * cx + F.R[1] * cy) + F.R[2] * cz) + F.t[0];
        const float py = ((F.R[3] * cx + F.R[4] * cy) + F.R[5] * cz) + F.t[1];
        const float pz = ((F.R[6] * cx + F.R[7] * cy) + F.R[8] * cz) + F.t[2];
        const float rz = __builtin_amdgcn_rcpf(pz);
        const float u = (px * vp.fx) * rz + vp.cx, v = (py * vp.fy) * rz + vp.cy;
        zmin = fminf(zmin, pz);
        umin = fminf(umin, u); umax = fmaxf(umax, u);
        vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
    }
    // corners at or behind the camera: the footprint is unbounded, keep
    if (!(zmin > 1e-3f)) return true;
    umin = fmaxf(umin - 2.0f, 0.0f); umax = fminf(umax + 2.0f, vp.fcols - 1.0f);
    vmin = fmaxf(vmin - 2.0f, 0.0f); vmax = fminf(vmax + 2.0f, vp.frows - 1.0f);
    if (!(umin <= umax && vmin <= vmax)) return false;  // footprint leaves the image
    const int iu0 = (int)umin, iu1 = (int)umax, iv0 = (int)vmin, iv1 = (int)vmax;
    const float margin = 1e-5f * (fabsf(zmin) + vp.trunc) + F.zmargin;
    const float need = zmin - vp.trunc - margin;
    // 32x32 blocks first (a superset of the footprint's 8x8 blocks): a coarse
    // maximum below `need` settles the test with a few loads
    if (!bound_rect_reaches(F.bound1, vp.b1x, iu0 / kBound1, iu1 / kBound1, iv0 / kBound1,
                            iv1 / kBound1, need))
        return false;
    const int bx0 = iu0 / kBound0, bx1 = iu1 / kBound0, by0 = iv0 / kBound0, by1 = iv1 / kBound0;
    if ((bx1 - bx0 + 1) * (by1 - by0 + 1) > 96) return true;
    return bound_rect_reaches(F.bound0, vp.b0x, bx0, bx1, by0, by1, need);
}

// Per 32x32 pixel block of frame blockIdx.z: the largest valid depth of each
// 8x8 sub-block (bound0) and of the block (bound1); 0 when none is valid.
// Valid depths are positive floats, so their bit patterns order like ints.
// Also clears the cull list counter for the launches that follow.
__global__ __launch_bounds__(256) void depth_bounds_kernel(const float* __restrict__ depth0,
                                                           size_t frame_stride, int pitch, int cols,
                                                           int rows, float maxd, float* bounds,
                                                           int b0x, int b0y, int b1x, int b1y,
                                                           unsigned* list_count) {
    __shared__ int s0[16];
    __shared__ int s1;
    const int t = threadIdx.x;
    if (t < 16) s0[t] = 0;
    if (t == 0) s1 = 0;
    if (t == 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) *list_count = 0;
    __syncthreads();
    const float* depth = depth0 + frame_stride * blockIdx.z;
    const int r = t >> 3, c4 = (t & 7) * 4;
    const int y = blockIdx.y * kBound1 + r;
    const int xb = blockIdx.x * kBound1 + c4;
    int m = 0;
    if (y < rows) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = xb + k;
            if (x < cols) {
                const float d = depth[y * pitch + x];
                if (d > 0.0f && d <= maxd) m = max(m, __float_as_int(d));
            }
        }
    }
    if (m) {
        atomicMax(&s0[(r >> 3) * 4 + (c4 >> 3)], m);
        atomicMax(&s1, m);
    }
    __syncthreads();
    float* out0 = bounds + (size_t)blockIdx.z * (b0x * b0y + b1x * b1y);
    float* out1 = out0 + b0x * b0y;
    if (t < 16) {
        const int bx = blockIdx.x * 4 + (t & 3), by = blockIdx.y * 4 + (t >> 2);
        if (bx < b0x && by < b0y) out0[by * b0x + bx] = __int_as_float(s0[t]);
    }
    if (t == 0) out1[blockIdx.y * b1x + blockIdx.x] = __int_as_float(s1);
}

// Frames of `nframes` whose frustum and depth bound reach wave-tile `id`.
template <int MAXF>
__device__ __forceinline__ unsigned tile_frames(const VolumeParams& vp, const int nframes,
                                                const FrameBlock<MAXF>& fb, uint32_t id) {
    int x0, y0, z0;
    tile_origin(vp, id, x0, y0, z0);
    float lo[3], hi[3];
    lo[0] = voxel_centre(x0, vp.vs[0], vp.off[0]);
    lo[1] = voxel_centre(y0, vp.vs[1], vp.off[1]);
    lo[2] = voxel_centre(z0, vp.vs[2], vp.off[2]);
    hi[0] = voxel_centre(min(x0 + kTileX - 1, vp.gx - 1), vp.vs[0], vp.off[0]);
    hi[1] = voxel_centre(min(y0 + kTileY - 1, vp.gy - 1), vp.vs[1], vp.off[1]);
    hi[2] = voxel_centre(min(z0 + kTileZ - 1, vp.gz - 1), vp.vs[2], vp.off[2]);
    unsigned live = 0;
    for (int f = 0; f < nframes; ++f) {
        bool in = true;
#pragma unroll
        for (int k = 0; k < 6; ++k) in = in && (plane_max(fb.f[f].plane[k], lo, hi) >= 0.0f);
        if (in && depth_reaches(vp, fb.f[f], lo, hi)) live |= 1u << f;
    }
    return live;
}

// Each thread tests kCullTilesPerThread wave-tiles; the workgroup's
// survivors are appended to `list` with one atomic per workgroup.  The list
// order varies from run to run; the result does not, since every voxel
// belongs to exactly one tile.
template <int MAXF>
__global__ __launch_bounds__(256) void cull_kernel(const VolumeParams vp, const int nframes,
                                                   const FrameBlock<MAXF> fb,
                                                   ListEntry* __restrict__ list,
                                                   unsigned* __restrict__ list_count) {
    __shared__ unsigned s_wave[kThreads / 64];
    __shared__ unsigned s_base;
    const uint32_t ntiles = (uint32_t)vp.tiles_x * vp.tiles_y * vp.tiles_z;
    const uint32_t first = blockIdx.x * (256u * kCullTilesPerThread) + threadIdx.x;
    unsigned live[kCullTilesPerThread];
    unsigned n = 0;
#pragma unroll
    for (int k = 0; k < kCullTilesPerThread; ++k) {
        const uint32_t id = first + 256u * k;
        live[k] = id < ntiles ? tile_frames(vp, nframes, fb, id) : 0u;
        n += live[k] != 0;
    }
    // exclusive prefix of n over the workgroup: within the wave, then waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    unsigned wave_base = 0, total = 0;
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i) {
        wave_base += i < wave ? s_wave[i] : 0u;
        total += s_wave[i];
    }
    if (threadIdx.x == 0) s_base = total ? atomicAdd(list_count, total) : 0u;
    __syncthreads();
    unsigned slot = s_base + wave_base + incl - n;
#pragma unroll
    for (int k = 0; k < kCullTilesPerThread; ++k)
        if (live[k]) list[slot++] = ListEntry{first + 256u * k, live[k]};
}

// Per-voxel work of one wave-tile.  Lane l owns voxels x..x+3 (x = x0 +
// 4*(l % kLanesX)) of row y = y0 + l / kLanesX for z = z0 + r, r < kRowsPerLane.
// Projection is branch-free so the 4*kRowsPerLane depth gathers of a frame
// issue back to back; D and W of a row are read only if one of its voxels
// passes, and written once after all frames.
template <int MAXF, bool VEC, bool COUNT>
__device__ __forceinline__ void integrate_tile(float* __restrict__ D, float* __restrict__ W,
                                               const VolumeParams& vp, const int nframes,
                                               const FrameBlock<MAXF>& fb, const unsigned live,
                                               const int x0, const int y0, const int z0,
                                               unsigned& cnt) {
    constexpr int NR = kRowsPerLane, NV = 4 * kRowsPerLane;
    const int lane = threadIdx.x & 63;
    const int x = x0 + 4 * (lane % kLanesX);
    const int y = y0 + lane / kLanesX;
    if (x >= vp.gx || y >= vp.gy) return;
    const int nvox = VEC ? 4 : min(4, vp.gx - x);
    int base[NR];
    bool row_ok[NR];
    float czw[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int z = z0 + r;
        row_ok[r] = z < vp.gz;
        base[r] = x + vp.gx * (y + vp.gy * min(z, vp.gz - 1));
        czw[r] = voxel_centre(z, vp.vs[2], vp.off[2]);
    }
    const float cyw = voxel_centre(y, vp.vs[1], vp.off[1]);
    float cxw[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) cxw[j] = voxel_centre(x + j, vp.vs[0], vp.off[0]);

    float d[NV], w[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) d[j] = w[j] = 0.0f;
    bool loaded[NR], dirty[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) loaded[r] = dirty[r] = false;

    for (int f = 0; f < nframes; ++f) {
        if (!((live >> f) & 1u)) continue;
        const FrameParams& F = fb.f[f];
        const float rxy = F.R[1] * cyw, ryy = F.R[4] * cyw, rzy = F.R[7] * cyw;
        bool ok[NV];
        float pz[NV], dep[NV];
        int addr[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int r = j >> 2, i = j & 3;
            const float px = ((F.R[0] * cxw[i] + rxy) + F.R[2] * czw[r]) + F.t[0];
            const float py = ((F.R[3] * cxw[i] + ryy) + F.R[5] * czw[r]) + F.t[1];
            pz[j] = ((F.R[6] * cxw[i] + rzy) + F.R[8] * czw[r]) + F.t[2];
            int u, v;
            const bool in_z = pz[j] > 0.0f;
            const bool in_u = pixel_index(px, vp.fx, pz[j], vp.cx, vp.fcols, u);
            const bool in_v = pixel_index(py, vp.fy, pz[j], vp.cy, vp.frows, v);
            ok[j] = in_z && in_u && in_v && row_ok[r] && (VEC || i < nvox);
            addr[j] = ok[j] ? v * F.pitch + u : 0;
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) dep[j] = F.depth[addr[j]];
        bool pass[NV], any[NR];
        float tsdf[NV];
#pragma unroll
        for (int r = 0; r < NR; ++r) any[r] = false;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const float sdf = dep[j] - pz[j];
            pass[j] = ok[j] && !(dep[j] <= 0.0f || dep[j] > vp.maxd) && (sdf >= -vp.trunc);
            tsdf[j] = sdf > 0.0f ? fminf(sdf, vp.trunc) : sdf;
            any[j >> 2] = any[j >> 2] || pass[j];
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (any[r] && !loaded[r]) {
                if (VEC) {
                    const float4 d4 = *reinterpret_cast<const float4*>(D + base[r]);
                    const float4 w4 = *reinterpret_cast<const float4*>(W + base[r]);
                    d[4 * r + 0] = d4.x; d[4 * r + 1] = d4.y; d[4 * r + 2] = d4.z; d[4 * r + 3] = d4.w;
                    w[4 * r + 0] = w4.x; w[4 * r + 1] = w4.y; w[4 * r + 2] = w4.z; w[4 * r + 3] = w4.w;
                } else {
                    for (int i = 0; i < nvox; ++i) {
                        d[4 * r + i] = D[base[r] + i];
                        w[4 * r + i] = W[base[r] + i];
                    }
                }
                loaded[r] = true;
            }
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            if (!pass[j]) continue;
            // tsdf_volume.cu:178-186 with the weight cap
            const float nw = w[j] + 1.0f;
            d[j] = ((d[j] * w[j]) + tsdf[j]) / nw;
            w[j] = nw < vp.maxw ? nw : vp.maxw;
            dirty[j >> 2] = true;
            if (COUNT) ++cnt;
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (!dirty[r]) continue;
        if (VEC) {
            *reinterpret_cast<float4*>(D + base[r]) =
                make_float4(d[4 * r + 0], d[4 * r + 1], d[4 * r + 2], d[4 * r + 3]);
            *reinterpret_cast<float4*>(W + base[r]) =
                make_float4(w[4 * r + 0], w[4 * r + 1], w[4 * r + 2], w[4 * r + 3]);
        } else {
            for (int i = 0; i < nvox; ++i) {
                D[base[r] + i] = d[4 * r + i];
                W[base[r] + i] = w[4 * r + i];
            }
        }
    }
}

// Each wavefront takes list entries i = wave, wave + waves, ...; the grid is
// sized to the resident wave slots, so every wave-tile costs about the same
// and the device stays full until the list is done.
template <int MAXF, bool VEC, bool COUNT>
__global__ __launch_bounds__(kThreads) void integrate_list_kernel(
    float* __restrict__ D, float* __restrict__ W, const VolumeParams vp, const int nframes,
    const FrameBlock<MAXF> fb, const ListEntry* __restrict__ list,
    const unsigned* __restrict__ list_count, unsigned long long* counter) {
    const unsigned n = __builtin_amdgcn_readfirstlane(*list_count);
    const unsigned waves = gridDim.x * (kThreads / 64);
    unsigned cnt = 0;
    for (unsigned i = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t tile = __builtin_amdgcn_readfirstlane(list[i].tile);
        const uint32_t frames = __builtin_amdgcn_readfirstlane(list[i].frames);
        int x0, y0, z0;
        tile_origin(vp, tile, x0, y0, z0);
        integrate_tile<MAXF, VEC, COUNT>(D, W, vp, nframes, fb, frames, x0, y0, z0, cnt);
    }
    if (COUNT && cnt) atomicAdd(counter, (unsigned long long)cnt);
}

// D <- value_d, W <- 0 (tsdf_volume.cu:21-26,230-238).
__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ D, float* __restrict__ W,
                                                   int n, float value_d) {
    const int i = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (i + 3 < n) {
        *reinterpret_cast<float4*>(D + i) = make_float4(value_d, value_d, value_d, value_d);
        *reinterpret_cast<float4*>(W + i) = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        for (int k = i; k < n; ++k) {
            D[k] = value_d;
            W[k] = 0.0f;
        }
    }
}

static int launch_fill(recon_tsdf* h, hipStream_t s) {
    const int n = (int)h->info.num_voxels;
    const int blocks = (n + 4 * 256 - 1) / (4 * 256);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, h->D, h->W, n,
                       h->info.truncation);
    RECON_HIP_TRY(hipGetLastError(), "fill_kernel launch");
    return RECON_OK;
}

static int check_depth(const recon_depth_view* depth, const char* fn) {
    if (!depth) return fail(RECON_ERR_NULL, "%s: depth view is NULL", fn);
    if (!depth->data) return fail(RECON_ERR_NULL, "%s: depth data is NULL", fn);
    if (depth->rows <= 0 || depth->cols <= 0)
        return fail(RECON_ERR_SIZE, "%s: depth size %dx%d must be positive", fn, depth->cols,
                    depth->rows);
    if (depth->step < 4 * (size_t)depth->cols || depth->step % 4 != 0)
        return fail(RECON_ERR_SIZE, "%s: depth step %zu must be a multiple of 4 and >= 4*cols", fn,
                    depth->step);
    if ((depth->step / 4) * (size_t)depth->rows >= (size_t(1) << 31))
        return fail(RECON_ERR_RANGE, "%s: depth image exceeds 2^31 floats", fn);
    return RECON_OK;
}

// Grows the handle's per-call scratch (depth bounds of kMaxFusedFrames frames,
// the cull list and its counter) to fit `vp`.  A buffer being replaced may
// still be read by queued kernels, so the stream is drained first.
static int ensure_scratch(recon_tsdf* h, const VolumeParams& vp, hipStream_t s) {
    const size_t per = (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
    const size_t need = per * kMaxFusedFrames;
    if (h->bounds_floats < need) {
        if (h->bounds) {
            RECON_HIP_TRY(hipStreamSynchronize(s), "scratch resize synchronize");
            (void)hipFree(h->bounds);
        }
        h->bounds = nullptr;
        h->bounds_floats = 0;
        RECON_HIP_TRY(hipMalloc((void**)&h->bounds, need * sizeof(float)), "depth bounds scratch");
        h->bounds_floats = need;
    }
    if (!h->list) {
        const size_t ntiles = (size_t)vp.tiles_x * vp.tiles_y * vp.tiles_z;
        RECON_HIP_TRY(hipMalloc(&h->list, ntiles * sizeof(ListEntry) + 256), "cull list");
        h->list_count = reinterpret_cast<unsigned*>(static_cast<char*>(h->list) +
                                                    ntiles * sizeof(ListEntry));
    }
    return RECON_OK;
}

// The three launches of one integrate sweep over `nframes` frames (frame f at
// depth0 + f * stride_floats).
template <int MAXF>
static int launch_integrate(recon_tsdf* h, const VolumeParams& vp, const float* depth0,
                            size_t stride_floats, int pitch, int nframes, FrameBlock<MAXF>& fb,
                            hipStream_t s) {
    int rc = ensure_scratch(h, vp, s);
    if (rc) return rc;
    const size_t per = (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
    for (int f = 0; f < nframes; ++f) {
        fb.f[f].bound0 = h->bounds + per * f;
        fb.f[f].bound1 = fb.f[f].bound0 + (size_t)vp.b0x * vp.b0y;
    }
    hipLaunchKernelGGL(depth_bounds_kernel, dim3(vp.b1x, vp.b1y, nframes), dim3(256), 0, s, depth0,
                       stride_floats, pitch, vp.cols, vp.rows, vp.maxd, h->bounds, vp.b0x, vp.b0y,
                       vp.b1x, vp.b1y, h->list_count);
    RECON_HIP_TRY(hipGetLastError(), "depth_bounds_kernel launch");
    const int ntiles = vp.tiles_x * vp.tiles_y * vp.tiles_z;
    ListEntry* list = static_cast<ListEntry*>(h->list);
    const int per_block = 256 * kCullTilesPerThread;
    hipLaunchKernelGGL((cull_kernel<MAXF>), dim3((ntiles + per_block - 1) / per_block), dim3(256), 0,
                       s, vp, nframes, fb, list, h->list_count);
    RECON_HIP_TRY(hipGetLastError(), "cull_kernel launch");
    const bool vec = (vp.gx % 4) == 0;
    unsigned long long* ctr = h->update_counter;
    const dim3 grid(h->num_cus * h->integrate_blocks_per_cu);
    if (ctr) {
        if (vec)
            hipLaunchKernelGGL((integrate_list_kernel<MAXF, true, true>), grid, dim3(kThreads), 0,
                               s, h->D, h->W, vp, nframes, fb, list, h->list_count, ctr);
        else
            hipLaunchKernelGGL((integrate_list_kernel<MAXF, false, true>), grid, dim3(kThreads), 0,
                               s, h->D, h->W, vp, nframes, fb, list, h->list_count, ctr);
    } else {
        if (vec)
            hipLaunchKernelGGL((integrate_list_kernel<MAXF, true, false>), grid, dim3(kThreads), 0,
                               s, h->D, h->W, vp, nframes, fb, list, h->list_count, ctr);
        else
            hipLaunchKernelGGL((integrate_list_kernel<MAXF, false, false>), grid, dim3(kThreads), 0,
                               s, h->D, h->W, vp, nframes, fb, list, h->list_count, ctr);
    }
    RECON_HIP_TRY(hipGetLastError(), "integrate_list_kernel launch");
    return RECON_OK;
}

static VolumeParams volume_params(const recon_tsdf* h, int cols, int rows) {
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

// Resident workgroups per CU of the integrate kernel (all instantiations
// share the register budget of the fused one, the largest).
static int integrate_occupancy() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, integrate_list_kernel<kMaxFusedFrames, true, false>, kThreads, 0) != hipSuccess ||
        blocks < 1)
        blocks = 4;
    int b1 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, integrate_list_kernel<1, true, false>,
                                                     kThreads, 0) == hipSuccess &&
        b1 >= 1)
        blocks = std::min(blocks, b1);
    return blocks;
}

}  // namespace recon

using namespace recon;

extern "C" int recon_tsdf_create(const recon_tsdf_config* cfg, recon_tsdf** out) {
    clear_error();
    if (!out) return fail(RECON_ERR_NULL, "recon_tsdf_create: out is NULL");
    *out = nullptr;
    recon_tsdf_info info;
    int rc = describe(cfg, &info);
    if (rc) return rc;
    rc = require_gfx950();
    if (rc) return rc;
    int device = 0, num_cus = 0;
    RECON_HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    RECON_HIP_TRY(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, device),
                  "hipDeviceGetAttribute");
    recon_tsdf* h = new recon_tsdf();
    h->device = device;
    h->num_cus = num_cus;
    h->integrate_blocks_per_cu = integrate_occupancy();
    h->cfg = *cfg;
    h->info = info;
    VolumeParams& vp = h->vp;
    std::memset(&vp, 0, sizeof(vp));
    vp.gx = info.grid[0];
    vp.gy = info.grid[1];
    vp.gz = info.grid[2];
    vp.tiles_x = (vp.gx + kTileX - 1) / kTileX;
    vp.tiles_y = (vp.gy + kTileY - 1) / kTileY;
    vp.tiles_z = (vp.gz + kTileZ - 1) / kTileZ;
    for (int i = 0; i < 3; ++i) {
        vp.vs[i] = info.voxel_size[i];
        vp.off[i] = info.global_offset[i];
    }
    vp.fx = info.intrinsics[0];
    vp.fy = info.intrinsics[1];
    vp.cx = info.intrinsics[2];
    vp.cy = info.intrinsics[3];
    vp.trunc = info.truncation;
    vp.maxw = info.max_weight;
    vp.maxd = info.max_depth;
    const size_t bytes = sizeof(float) * (size_t)info.num_voxels;
    hipError_t e = hipMalloc((void**)&h->D, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&h->W, bytes);
    if (e != hipSuccess) {
        rc = hip_fail(e, "recon_tsdf_create: hipMalloc");
        recon_tsdf_destroy(h);
        return rc;
    }
    rc = launch_fill(h, nullptr);
    if (rc == RECON_OK) {
        e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = hip_fail(e, "recon_tsdf_create: synchronize");
    }
    if (rc) {
        recon_tsdf_destroy(h);
        return rc;
    }
    *out = h;
    return RECON_OK;
}

extern "C" void recon_tsdf_destroy(recon_tsdf* h) {
    if (!h) return;
    if (h->D) (void)hipFree(h->D);
    if (h->W) (void)hipFree(h->W);
    if (h->mc_scratch) (void)hipFree(h->mc_scratch);
    if (h->bounds) (void)hipFree(h->bounds);
    if (h->list) (void)hipFree(h->list);
    delete h;
}

extern "C" int recon_tsdf_reset(recon_tsdf* h, recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_reset: handle is NULL");
    return launch_fill(h, (hipStream_t)stream);
}

extern "C" int recon_tsdf_get_info(const recon_tsdf* h, recon_tsdf_info* info) {
    clear_error();
    if (!h || !info) return fail(RECON_ERR_NULL, "recon_tsdf_get_info: null argument");
    *info = h->info;
    return RECON_OK;
}

extern "C" float* recon_tsdf_distances(recon_tsdf* h) { return h ? h->D : nullptr; }
extern "C" float* recon_tsdf_weights(recon_tsdf* h) { return h ? h->W : nullptr; }

extern "C" int recon_tsdf_download(recon_tsdf* h, float* host_d, float* host_w,
                                   recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_download: handle is NULL");
    const size_t bytes = sizeof(float) * (size_t)h->info.num_voxels;
    hipStream_t s = (hipStream_t)stream;
    if (host_d) RECON_HIP_TRY(hipMemcpyAsync(host_d, h->D, bytes, hipMemcpyDeviceToHost, s), "download D");
    if (host_w) RECON_HIP_TRY(hipMemcpyAsync(host_w, h->W, bytes, hipMemcpyDeviceToHost, s), "download W");
    RECON_HIP_TRY(hipStreamSynchronize(s), "recon_tsdf_download: synchronize");
    return RECON_OK;
}

extern "C" int recon_tsdf_upload(recon_tsdf* h, const float* host_d, const float* host_w,
                                 recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_upload: handle is NULL");
    const size_t bytes = sizeof(float) * (size_t)h->info.num_voxels;
    hipStream_t s = (hipStream_t)stream;
    if (host_d) RECON_HIP_TRY(hipMemcpyAsync(h->D, host_d, bytes, hipMemcpyHostToDevice, s), "upload D");
    if (host_w) RECON_HIP_TRY(hipMemcpyAsync(h->W, host_w, bytes, hipMemcpyHostToDevice, s), "upload W");
    RECON_HIP_TRY(hipStreamSynchronize(s), "recon_tsdf_upload: synchronize");
    return RECON_OK;
}

extern "C" int recon_tsdf_set_update_counter(recon_tsdf* h, unsigned long long* counter) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_set_update_counter: handle is NULL");
    h->update_counter = counter;
    return RECON_OK;
}

// TSDFVolume::integrate (tsdf_volume.cu:257-290): the reference rejects a
// null colour image; colour is never read, so here it is optional.
extern "C" int recon_tsdf_integrate(recon_tsdf* h, const recon_depth_view* depth,
                                    const uint8_t* rgb, size_t rgb_step,
                                    const double camera_pose[7], recon_stream_t stream) {
    (void)rgb;
    (void)rgb_step;
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_integrate: handle is NULL");
    int rc = check_depth(depth, "recon_tsdf_integrate");
    if (rc) return rc;
    if (!camera_pose) return fail(RECON_ERR_NULL, "recon_tsdf_integrate: pose is NULL");
    if (!pose_is_finite(camera_pose))
        return fail(RECON_ERR_CONFIG, "recon_tsdf_integrate: pose has a non-finite entry");
    const VolumeParams vp = volume_params(h, depth->cols, depth->rows);
    FrameBlock<1> fb;
    const int pitch = (int32_t)(depth->step / 4);
    make_frame_params(vp, camera_pose, depth->data, pitch, &fb.f[0]);
    return launch_integrate<1>(h, vp, depth->data, 0, pitch, 1, fb, (hipStream_t)stream);
}

extern "C" int recon_tsdf_integrate_frames(recon_tsdf* h, const recon_depth_view* depth0,
                                           size_t frame_stride_bytes, int32_t num_frames,
                                           const double* camera_poses, recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_integrate_frames: handle is NULL");
    int rc = check_depth(depth0, "recon_tsdf_integrate_frames");
    if (rc) return rc;
    if (!camera_poses) return fail(RECON_ERR_NULL, "recon_tsdf_integrate_frames: poses is NULL");
    if (num_frames < 0) return fail(RECON_ERR_SIZE, "recon_tsdf_integrate_frames: num_frames < 0");
    if (num_frames > 1 && frame_stride_bytes % 4 != 0)
        return fail(RECON_ERR_SIZE, "recon_tsdf_integrate_frames: frame stride not a multiple of 4");
    for (int f = 0; f < num_frames; ++f)
        if (!pose_is_finite(camera_poses + 7 * f))
            return fail(RECON_ERR_CONFIG, "recon_tsdf_integrate_frames: pose %d non-finite", f);
    const VolumeParams vp = volume_params(h, depth0->cols, depth0->rows);
    const char* base = reinterpret_cast<const char*>(depth0->data);
    for (int f0 = 0; f0 < num_frames; f0 += kMaxFusedFrames) {
        const int n = std::min(kMaxFusedFrames, num_frames - f0);
        FrameBlock<kMaxFusedFrames> fb;
        std::memset(&fb, 0, sizeof(fb));
        for (int k = 0; k < n; ++k) {
            const float* dptr = reinterpret_cast<const float*>(base + (size_t)(f0 + k) * frame_stride_bytes);
            make_frame_params(vp, camera_poses + 7 * (f0 + k), dptr, (int32_t)(depth0->step / 4),
                              &fb.f[k]);
        }
        rc = launch_integrate<kMaxFusedFrames>(
            h, vp, reinterpret_cast<const float*>(base + (size_t)f0 * frame_stride_bytes),
            frame_stride_bytes / 4, (int32_t)(depth0->step / 4), n, fb, (hipStream_t)stream);
        if (rc) return rc;
    }
    return RECON_OK;
}
