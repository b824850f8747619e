// recon_common.hpp -- host-side state, error reporting and the float32 voxel
// math shared by the gfx950 kernels of the TSDF fusion path.
//
// Arithmetic contract (identical in every kernel and in the CPU oracle):
// IEEE float32, one rounding per operation in the order written, no FMA
// contraction (translation units carry `#pragma clang fp contract(off)` and
// are built with -ffp-contract=off), correctly rounded division, denormals
// kept.  The operation order is the reference's:
//   voxel centre     ((v + 0.5f) * vs) + off                tsdf_volume.cu:117-119
//   pose -> matrix   SE3(float4 quat, float3 t)              se3.cuh:42-57,124-140
//   world -> camera  ((r0*x + r1*y) + r2*z) + t              se3.cuh:82-100
//   camera -> pixel  (int)(((x * fx) / z) + cx)              camera_util.cu:7-19
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "recon/recon.h"

namespace recon {

// ---- error reporting -------------------------------------------------------
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
int hip_fail(hipError_t e, const char* what);

#define RECON_HIP_TRY(expr, what)                         \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return ::recon::hip_fail(_e, what); \
    } while (0)

// Checks that the current device is a gfx950 part; caches the answer per
// device.  Returns RECON_OK or RECON_ERR_DEVICE.
int require_gfx950();

// ---- pose ------------------------------------------------------------------
// {qw,qx,qy,qz,tx,ty,tz} (double) -> float32 row-major R and t, narrowing the
// way tsdf_volume.cu:271-272 does before building SE3<float>.
void pose_to_matrix_f32(const double pose[7], float R[9], float t[3]);
bool pose_is_finite(const double pose[7]);

// ---- per-frame kernel parameters ------------------------------------------
// Conservative tile culling: a voxel can only be updated when its camera
// point p satisfies a_k . p + b_k > 0 for six half-spaces (z > 0, the four
// image borders u > -1, u < cols, v > -1, v < rows, and z < max_depth +
// truncation).  In world coordinates each is n_k . c + k_k > 0; `plane`
// holds (n_k, k_k + margin_k), margin_k = 1e-3 of the half-space's magnitude
// over the volume, three orders above float32 rounding of the per-voxel
// test, so a tile whose voxel-centre box has max_c(n.c + k + margin) < 0
// holds no voxel that passes.
//
// Depth-bound culling: `bound0` / `bound1` hold, per 8x8 / 32x32 pixel block
// of the frame, the largest valid depth (0 < d <= max_depth) in the block, or
// 0 when the block has none.  A voxel can only be updated through a pixel of
// its tile's image footprint whose depth is >= (voxel camera z) - truncation,
// so a tile whose footprint maximum is below (tile's least camera z) -
// truncation - margin holds no voxel that passes.
struct FrameParams {
    float R[9];
    float t[3];
    float plane[6][4];
    const float* depth;
    const float* bound0;
    const float* bound1;
    int32_t pitch;   // floats per depth row
    float zmargin;   // bound on float32 error of a camera z: 1e-6 * (|t| + max|c|)
    const uint8_t* rgb = nullptr;  // colour image (rgb triples) of the colour integrate
    int32_t rgb_pitch = 0;         // bytes per colour row
};

constexpr int kBound0 = 8, kBound1 = 32;  // pixel block edges of the two bound levels
// Segments of a segmented cull list (one count word each after the entries);
// the consumer keeps their offsets in 16 KB of LDS.
constexpr int kMaxSegments = 4096;

struct VolumeParams {
    float* D;
    float* W;
    int32_t gx, gy, gz;
    int32_t tiles_x, tiles_y, tiles_z;
    float vs[3];
    float off[3];
    float fx, fy, cx, cy;
    float trunc, maxw, maxd;
    int32_t cols, rows;
    float fcols, frows;
    int32_t b0x, b0y, b1x, b1y;  // bound-level block counts
    float* C = nullptr;          // colour volume, 4 floats per voxel (r, g, b, weight)
    float noise[3] = {0.f, 0.f, 0.f};  // WEIGHT_BY_VARIANCE: sigma(d) = a + b (d - c)^2
    float wref = 0.f;                  // sigma(1 m)^2
};

// Observation weight of depth sample d under WEIGHT_BY_VARIANCE:
// sigma(1 m)^2 / sigma(d)^2 with sigma(d) = a + (b * ((d - c) * (d - c))).
__host__ __device__ inline float variance_weight(const float noise[3], float wref, float d) {
    const float e = d - noise[2];
    const float sig = noise[0] + (noise[1] * (e * e));
    return wref / (sig * sig);
}

constexpr int kMaxFusedFrames = 16;

// Fills FrameParams for one frame of a volume.
void make_frame_params(const VolumeParams& vp, const double pose[7], const float* depth,
                       int32_t pitch_floats, FrameParams* fp);

// The culling half-spaces and z margin of a frame from its R and t (float64
// arithmetic, the same on host and device).  Half-space k in camera space is
// a_k . p + b_k > 0; in world space n = R^T a_k, k = a_k . t + b_k, widened
// by 1e-3 of the largest magnitude the terms reach over the volume.
__host__ __device__ inline void frame_planes(const VolumeParams& vp, FrameParams* fp) {
    const double fx = vp.fx, fy = vp.fy, cx = vp.cx, cy = vp.cy;
    const double a[6][3] = {{0, 0, 1},
                            {fx, 0, cx + 1.0},
                            {-fx, 0, (double)vp.cols - cx},
                            {0, fy, cy + 1.0},
                            {0, -fy, (double)vp.rows - cy},
                            {0, 0, -1}};
    const double b[6] = {0, 0, 0, 0, 0, (double)vp.maxd + (double)vp.trunc};
    const int g[3] = {vp.gx, vp.gy, vp.gz};
    double cmax[3];
    for (int i = 0; i < 3; ++i) {
        const double lo = vp.off[i], hi = (double)vp.off[i] + (double)g[i] * vp.vs[i];
        cmax[i] = ::fmax(::fabs(lo), ::fabs(hi)) + ::fabs((double)vp.vs[i]);
    }
    const double tabs = ::fabs((double)fp->t[0]) + ::fabs((double)fp->t[1]) + ::fabs((double)fp->t[2]);
    fp->zmargin = (float)(1e-6 * (tabs + cmax[0] + cmax[1] + cmax[2]) + 1e-6);
    for (int k = 0; k < 6; ++k) {
        double n[3];
        for (int i = 0; i < 3; ++i)  // n = R^T a
            n[i] = (double)fp->R[0 * 3 + i] * a[k][0] + (double)fp->R[1 * 3 + i] * a[k][1] +
                   (double)fp->R[2 * 3 + i] * a[k][2];
        const double kk = a[k][0] * fp->t[0] + a[k][1] * fp->t[1] + a[k][2] * fp->t[2] + b[k];
        const double a1 = ::fabs(a[k][0]) + ::fabs(a[k][1]) + ::fabs(a[k][2]);
        double scale = a1 * tabs + ::fabs(b[k]);
        for (int i = 0; i < 3; ++i) scale += ::fabs(n[i]) * cmax[i];
        const double margin = 1e-3 * scale + 1e-12;
        for (int i = 0; i < 3; ++i) fp->plane[k][i] = (float)n[i];
        fp->plane[k][3] = (float)(kk + margin);
    }
}

}  // namespace recon

// ---- the handle -------------------------------------------------------------
struct recon_tsdf {
    recon_tsdf_config cfg;
    recon_tsdf_info info;
    recon::VolumeParams vp;  // cols/rows filled per call
    float* D = nullptr;
    float* W = nullptr;
    int device = 0;
    int num_cus = 256;
    unsigned long long* update_counter = nullptr;
    unsigned long long* touch_counter = nullptr;
    // per-frame depth bounds, grown on demand
    float* bounds = nullptr;
    size_t bounds_floats = 0;
    void* list = nullptr;            // cull list: one entry per wave-tile, then
    unsigned* list_count = nullptr;  // kMaxSegments count words (inside `list`)
    // marching-cubes scratch, grown on demand
    void* mc_scratch = nullptr;
    size_t mc_scratch_bytes = 0;
    // differentiable path scratch (frame table, depth bounds, partial sums)
    void* diff_scratch = nullptr;
    size_t diff_scratch_bytes = 0;
    uint32_t* owner = nullptr;  // per-voxel table of the adjoint's pair merge
    bool owner_dirty = false;   // set while a merge is being queued
    float* C = nullptr;         // colour volume (recon_tsdf_enable_color), 16 B per voxel
};

namespace recon {

// Validates a configuration and derives voxel size / truncation / max depth
// (tsdf_volume.cu:194-199).  Host only, no HIP calls.
int describe(const recon_tsdf_config* cfg, recon_tsdf_info* info);

// Allocates the volume's cull list (`ntiles` entries of `entry_bytes`, then
// the count words) on first use.
int ensure_cull_list(recon_tsdf* h, size_t ntiles, size_t entry_bytes);

// Device scratch bytes marching cubes needs for `grid`.
size_t mc_scratch_need(const int32_t grid[3]);

// Marching cubes over a dense field; `scratch` / `scratch_bytes` may point
// at a reusable buffer owned by the caller, grown by the callee when
// `may_allocate`; otherwise a buffer smaller than mc_scratch_need fails.
int mc_run(const float* field, const int32_t grid[3], const float vs[3], const float off[3],
           float* verts, int32_t* tris, int64_t capacity, int64_t* total, bool count_only,
           void** scratch, size_t* scratch_bytes, hipStream_t stream, bool may_allocate = true);

}  // namespace recon

// ---- device math --------------------------------------------------------------
namespace recon {

__device__ __forceinline__ float voxel_centre(int v, float vs, float off) {
    return (((float)v + 0.5f) * vs) + off;
}

// Pixel index along one axis of camera point (pa, pz): s = ((pa*f)/pz) + c.
// Valid iff -1 < s < size, which equals 0 <= (int)s < size for finite s.
__device__ __forceinline__ bool pixel_index(float pa, float f, float pz, float c, float size,
                                            int& out) {
    const float s = ((pa * f) / pz) + c;
    if (!(s > -1.0f && s < size)) return false;
    out = (int)s;
    return true;
}

// The correctly rounded quotient a/b from y = RN(1/b) (Markstein's
// correction, P. Markstein, IBM J. Res. Dev. 34(1), 1990): q = RN(a*y),
// r = a - b*q (exact, one FMA), RN(q + r*y) = RN(a/b) whenever a/b is in
// the normal range (or zero: then the result is +0 where a = -0).
// tests/div_correction_check.c checks it against IEEE division for every
// float32 mantissa of a with the integer divisors 1..256, and for random
// a, b over the normal range.
__device__ __forceinline__ float div_by_rcp(float a, float b, float y) {
    const float q = a * y;
    return __builtin_fmaf(__builtin_fmaf(-q, b, a), y, q);
}

// pixel_index with rz = RN(1/pz) (1.0f / pz) shared by both axes: the same
// index and validity.  Outside the normal range (pz near 0, or an overflow)
// both quotients are infinite, NaN or below ulp(c) / 2, and both indices
// invalid or equal after + c and the truncation.
__device__ __forceinline__ bool pixel_index_rcp(float pa, float f, float pz, float rz, float c,
                                                float size, int& out) {
    const float s = div_by_rcp(pa * f, pz, rz) + c;
    if (!(s > -1.0f && s < size)) return false;
    out = (int)s;
    return true;
}


}  // namespace recon
