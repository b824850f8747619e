// tsdf_volume.hip -- TSDF volume lifecycle and the projective integrate
// kernel for gfx950.
//
// Replaces TSDFVolume (cu-Reconstruction modules/src/cuda/tsdf_volume.cu:
// 144-290).  The reference runs one thread per (y,z) voxel row looping over x
// (tsdf_volume.cu:151-163, uncoalesced: neighbouring lanes are a row apart)
// and re-reads a 28-byte voxel-centre table per voxel.  Here:
//   * a workgroup owns a 64 x 4 x 4 voxel tile; lane l covers 4 consecutive
//     x voxels of one row, so one wave moves four contiguous 256-byte row
//     segments of D and W per dwordx4 access;
//   * voxel centres are computed from the index (no table traffic);
//   * the tile is tested against six half-spaces of the camera frustum
//     before any voxel work, and D/W are only loaded by lanes holding a voxel
//     that passes every test of the frame (wavefront-level early-out);
//   * up to kMaxFusedFrames frames can be applied in one sweep: each voxel
//     sees the frames in order, so the result equals one sweep per frame.
#pragma clang fp contract(off)

#include "recon_common.hpp"

#include <cmath>
#include <cstring>

namespace recon {

constexpr int kTileX = 64, kTileY = 4, kTileZ = 4, kThreads = 256;

template <int MAXF>
struct FrameBlock {
    FrameParams f[MAXF];
};

// Max over the tile's voxel-centre box of n.c + k.
__device__ __forceinline__ float plane_max(const float* p, const float lo[3], const float hi[3]) {
    float g = p[3];
    g = g + fmaxf(p[0] * lo[0], p[0] * hi[0]);
    g = g + fmaxf(p[1] * lo[1], p[1] * hi[1]);
    g = g + fmaxf(p[2] * lo[2], p[2] * hi[2]);
    return g;
}

template <int MAXF, bool VEC, bool COUNT>
__global__ __launch_bounds__(kThreads) void integrate_kernel(float* __restrict__ D,
                                                             float* __restrict__ W,
                                                             const VolumeParams vp, const int nframes,
                                                             const FrameBlock<MAXF> fb,
                                                             unsigned long long* counter) {
    const int tile = blockIdx.x;
    const int tx = tile % vp.tiles_x;
    const int tyz = tile / vp.tiles_x;
    const int ty = tyz % vp.tiles_y;
    const int tz = tyz / vp.tiles_y;
    const int x0 = tx * kTileX, y0 = ty * kTileY, z0 = tz * kTileZ;

    // voxel-centre box of the tile
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
        live |= (in ? 1u : 0u) << f;
    }
    if (live == 0) return;

    const int tid = threadIdx.x;
    const int x = x0 + 4 * (tid & 15);
    const int row = tid >> 4;
    const int y = y0 + (row & 3);
    const int z = z0 + (row >> 2);
    if (x >= vp.gx || y >= vp.gy || z >= vp.gz) return;
    const int nvox = VEC ? 4 : min(4, vp.gx - x);
    const int base = x + vp.gx * (y + vp.gy * z);

    const float cyw = voxel_centre(y, vp.vs[1], vp.off[1]);
    const float czw = voxel_centre(z, vp.vs[2], vp.off[2]);
    float cxw[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) cxw[j] = voxel_centre(x + j, vp.vs[0], vp.off[0]);

    float d[4] = {0.f, 0.f, 0.f, 0.f}, w[4] = {0.f, 0.f, 0.f, 0.f};
    bool loaded = false, dirty = false;
    unsigned cnt = 0;

    for (int f = 0; f < nframes; ++f) {
        if (!((live >> f) & 1u)) continue;
        const FrameParams& F = fb.f[f];
        const float rxy = F.R[1] * cyw, rxz = F.R[2] * czw;
        const float ryy = F.R[4] * cyw, ryz = F.R[5] * czw;
        const float rzy = F.R[7] * cyw, rzz = F.R[8] * czw;
        bool pass[4];
        float tsdf[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            pass[j] = false;
            tsdf[j] = 0.0f;
            if (!VEC && j >= nvox) continue;
            const float px = ((F.R[0] * cxw[j] + rxy) + rxz) + F.t[0];
            const float py = ((F.R[3] * cxw[j] + ryy) + ryz) + F.t[1];
            const float pz = ((F.R[6] * cxw[j] + rzy) + rzz) + F.t[2];
            if (!(pz > 0.0f)) continue;
            int u, v;
            if (!pixel_index(px, vp.fx, pz, vp.cx, vp.fcols, u)) continue;
            if (!pixel_index(py, vp.fy, pz, vp.cy, vp.frows, v)) continue;
            const float dep = F.depth[v * F.pitch + u];
            if (dep <= 0.0f || dep > vp.maxd) continue;
            const float sdf = dep - pz;
            if (!(sdf >= -vp.trunc)) continue;
            tsdf[j] = sdf > 0.0f ? fminf(sdf, vp.trunc) : sdf;
            pass[j] = true;
        }
        if (pass[0] || pass[1] || pass[2] || pass[3]) {
            if (!loaded) {
                if (VEC) {
                    const float4 d4 = *reinterpret_cast<const float4*>(D + base);
                    const float4 w4 = *reinterpret_cast<const float4*>(W + base);
                    d[0] = d4.x; d[1] = d4.y; d[2] = d4.z; d[3] = d4.w;
                    w[0] = w4.x; w[1] = w4.y; w[2] = w4.z; w[3] = w4.w;
                } else {
                    for (int j = 0; j < nvox; ++j) {
                        d[j] = D[base + j];
                        w[j] = W[base + j];
                    }
                }
                loaded = true;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (!pass[j]) continue;
                const float nw = w[j] + 1.0f;
                d[j] = ((d[j] * w[j]) + tsdf[j]) / nw;
                w[j] = nw < vp.maxw ? nw : vp.maxw;
                if (COUNT) ++cnt;
            }
            dirty = true;
        }
    }
    if (dirty) {
        if (VEC) {
            *reinterpret_cast<float4*>(D + base) = make_float4(d[0], d[1], d[2], d[3]);
            *reinterpret_cast<float4*>(W + base) = make_float4(w[0], w[1], w[2], w[3]);
        } else {
            for (int j = 0; j < nvox; ++j) {
                D[base + j] = d[j];
                W[base + j] = w[j];
            }
        }
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
    h->version++;
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

template <int MAXF>
static int launch_integrate(recon_tsdf* h, const VolumeParams& vp, int nframes,
                            const FrameBlock<MAXF>& fb, hipStream_t s) {
    const int tiles = vp.tiles_x * vp.tiles_y * vp.tiles_z;
    const bool vec = (vp.gx % 4) == 0;
    unsigned long long* ctr = h->update_counter;
    if (ctr) {
        if (vec)
            hipLaunchKernelGGL((integrate_kernel<MAXF, true, true>), dim3(tiles), dim3(kThreads), 0,
                               s, h->D, h->W, vp, nframes, fb, ctr);
        else
            hipLaunchKernelGGL((integrate_kernel<MAXF, false, true>), dim3(tiles), dim3(kThreads),
                               0, s, h->D, h->W, vp, nframes, fb, ctr);
    } else {
        if (vec)
            hipLaunchKernelGGL((integrate_kernel<MAXF, true, false>), dim3(tiles), dim3(kThreads),
                               0, s, h->D, h->W, vp, nframes, fb, ctr);
        else
            hipLaunchKernelGGL((integrate_kernel<MAXF, false, false>), dim3(tiles),
                               dim3(kThreads), 0, s, h->D, h->W, vp, nframes, fb, ctr);
    }
    RECON_HIP_TRY(hipGetLastError(), "integrate_kernel launch");
    h->version++;
    return RECON_OK;
}

static VolumeParams volume_params(const recon_tsdf* h, int cols, int rows) {
    VolumeParams vp = h->vp;
    vp.cols = cols;
    vp.rows = rows;
    vp.fcols = (float)cols;
    vp.frows = (float)rows;
    return vp;
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
    recon_tsdf* h = new recon_tsdf();
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
    RECON_HIP_TRY(hipGetDevice(&h->device), "hipGetDevice");
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
    h->version++;
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
    make_frame_params(vp, camera_pose, depth->data, (int32_t)(depth->step / 4), &fb.f[0]);
    return launch_integrate<1>(h, vp, 1, fb, (hipStream_t)stream);
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
        rc = launch_integrate<kMaxFusedFrames>(h, vp, n, fb, (hipStream_t)stream);
        if (rc) return rc;
    }
    return RECON_OK;
}
