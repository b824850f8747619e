// recon_common.hip -- error reporting, device check, configuration
// validation and per-frame parameter set-up for the TSDF fusion path.
#pragma clang fp contract(off)

#include "recon_common.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

namespace recon {

static thread_local std::string t_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_last_error = buf;
    return code;
}

void clear_error() { t_last_error.clear(); }

int hip_fail(hipError_t e, const char* what) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? RECON_ERR_OOM
                                                                                 : RECON_ERR_HIP;
    return fail(code, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

int require_gfx950() {
    static std::mutex mu;
    static int state[64] = {0};  // 0 unknown, 1 ok, -1 rejected
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(RECON_ERR_DEVICE, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(mu);
    if (state[dev] == 0) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, dev);
        if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
        state[dev] = (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) ? 1 : -1;
        if (state[dev] < 0)
            return fail(RECON_ERR_DEVICE, "device %d is %s; this library is built for gfx950", dev,
                        prop.gcnArchName);
    }
    if (state[dev] < 0) return fail(RECON_ERR_DEVICE, "device %d is not gfx950", dev);
    return RECON_OK;
}

bool pose_is_finite(const double pose[7]) {
    for (int i = 0; i < 7; ++i)
        if (!std::isfinite(pose[i])) return false;
    return true;
}

// se3.cuh:42-57,124-140 after the narrowing of tsdf_volume.cu:271-272.
void pose_to_matrix_f32(const double pose[7], float R[9], float t[3]) {
    const float qw = (float)pose[0], qx = (float)pose[1], qy = (float)pose[2], qz = (float)pose[3];
    const float x = 2.0f * qx, y = 2.0f * qy, z = 2.0f * qz;
    const float wx = x * qw, wy = y * qw, wz = z * qw;
    const float xx = x * qx, xy = y * qx, xz = z * qx;
    const float yy = y * qy, yz = z * qy, zz = z * qz;
    R[0] = 1.0f - (yy + zz);
    R[1] = xy - wz;
    R[2] = xz + wy;
    R[3] = xy + wz;
    R[4] = 1.0f - (xx + zz);
    R[5] = yz - wx;
    R[6] = xz - wy;
    R[7] = yz + wx;
    R[8] = 1.0f - (xx + yy);
    t[0] = (float)pose[4];
    t[1] = (float)pose[5];
    t[2] = (float)pose[6];
}

static bool finite_pos(float v) { return std::isfinite(v) && v > 0.0f; }

// TSDFVolume::TSDFVolume (tsdf_volume.cu:194-199): voxel_size = physical /
// grid (float3 / dim3, cudaMath.cuh:1225-1228), truncation 1.1*|voxel_size|.
int ensure_cull_list(recon_tsdf* h, size_t ntiles, size_t entry_bytes) {
    if (h->list) return RECON_OK;
    RECON_HIP_TRY(hipMalloc(&h->list, ntiles * entry_bytes + sizeof(unsigned) * kMaxSegments),
                  "cull list");
    h->list_count = reinterpret_cast<unsigned*>(static_cast<char*>(h->list) + ntiles * entry_bytes);
    return RECON_OK;
}

int describe(const recon_tsdf_config* cfg, recon_tsdf_info* info) {
    if (!cfg || !info) return fail(RECON_ERR_NULL, "recon_tsdf_describe: null argument");
    int64_t n = 1;
    for (int i = 0; i < 3; ++i) {
        if (cfg->voxel_grid_size[i] < 1)
            return fail(RECON_ERR_CONFIG, "voxel_grid_size[%d] = %d must be >= 1", i,
                        cfg->voxel_grid_size[i]);
        if (!finite_pos(cfg->physical_size[i]))
            return fail(RECON_ERR_CONFIG, "physical_size[%d] must be finite and > 0", i);
        if (!std::isfinite(cfg->global_offset[i]))
            return fail(RECON_ERR_CONFIG, "global_offset[%d] must be finite", i);
        n *= cfg->voxel_grid_size[i];
    }
    if (n >= (int64_t(1) << 31))
        return fail(RECON_ERR_RANGE, "grid of %lld voxels exceeds 2^31-1", (long long)n);
    if (!(std::isfinite(cfg->max_weight) && cfg->max_weight >= 1.0f))
        return fail(RECON_ERR_CONFIG, "max_weight must be finite and >= 1");
    if (!finite_pos(cfg->focal_x) || !finite_pos(cfg->focal_y))
        return fail(RECON_ERR_CONFIG, "focal lengths must be finite and > 0");
    if (!std::isfinite(cfg->cx) || !std::isfinite(cfg->cy))
        return fail(RECON_ERR_CONFIG, "principal point must be finite");
    if (!std::isfinite(cfg->truncation) || !std::isfinite(cfg->max_depth))
        return fail(RECON_ERR_CONFIG, "truncation / max_depth must be finite");
    std::memset(info, 0, sizeof(*info));
    for (int i = 0; i < 3; ++i) {
        info->grid[i] = cfg->voxel_grid_size[i];
        info->voxel_size[i] = cfg->physical_size[i] / (float)cfg->voxel_grid_size[i];
        info->global_offset[i] = cfg->global_offset[i];
    }
    const float* vs = info->voxel_size;
    const float len = std::sqrt(vs[0] * vs[0] + vs[1] * vs[1] + vs[2] * vs[2]);
    info->truncation = cfg->truncation > 0.0f ? cfg->truncation : 1.1f * len;
    info->max_depth = cfg->max_depth > 0.0f ? cfg->max_depth : 24.0f;
    info->max_weight = cfg->max_weight;
    info->intrinsics[0] = cfg->focal_x;
    info->intrinsics[1] = cfg->focal_y;
    info->intrinsics[2] = cfg->cx;
    info->intrinsics[3] = cfg->cy;
    info->num_voxels = n;
    if (!finite_pos(info->truncation)) return fail(RECON_ERR_CONFIG, "truncation must be > 0");
    if (cfg->weight_type != RECON_WEIGHT_BY_DEPTH && cfg->weight_type != RECON_WEIGHT_BY_VARIANCE)
        return fail(RECON_ERR_CONFIG, "weight_type %d is not a WEIGHT_TYPE", cfg->weight_type);
    info->weight_type = cfg->weight_type;
    const float* nm = cfg->noise_model;
    if (nm[0] == 0.0f && nm[1] == 0.0f && nm[2] == 0.0f) {
        // axial noise of a structured-light depth camera (Nguyen, Izadi, Lovell 2012)
        info->noise_model[0] = 0.0012f;
        info->noise_model[1] = 0.0019f;
        info->noise_model[2] = 0.4f;
    } else {
        if (!finite_pos(nm[0]) || !std::isfinite(nm[1]) || nm[1] < 0.0f || !std::isfinite(nm[2]))
            return fail(RECON_ERR_CONFIG, "noise_model needs a > 0, b >= 0, finite c");
        for (int i = 0; i < 3; ++i) info->noise_model[i] = nm[i];
    }
    return RECON_OK;
}

void make_frame_params(const VolumeParams& vp, const double pose[7], const float* depth,
                       int32_t pitch_floats, FrameParams* fp) {
    pose_to_matrix_f32(pose, fp->R, fp->t);
    fp->depth = depth;
    fp->bound0 = nullptr;
    fp->bound1 = nullptr;
    fp->pitch = pitch_floats;
    frame_planes(vp, fp);
}

}  // namespace recon

extern "C" const char* recon_last_error(void) { return recon::t_last_error.c_str(); }
extern "C" int recon_abi_version(void) { return RECON_ABI_VERSION; }
extern "C" int recon_tsdf_describe(const recon_tsdf_config* cfg, recon_tsdf_info* info) {
    recon::clear_error();
    return recon::describe(cfg, info);
}
