// render_points.hip -- z-buffer render of a world point cloud into a depth map
// and an rgb image for gfx950.
//
// Replaces cudaPointCloud::project_to_depth / project_to_color_depth and
// filter_image_if (cu-Reconstruction modules/src/cuda/cuda_point_cloud.cu:
// 229-353, 89-99).  The reference splats on the host, one point at a time in
// index order: a pixel keeps the smallest camera z, the first of equal ones,
// and that point's colour; pixels nobody wrote (still FLT_MAX) become 0.
//
// Here the order-dependent host loop becomes an order-free device reduction
// with the same result: every point forms the 64-bit key
//     (order-preserving bits of z) << 32 | point index
// and the pixel takes the atomic minimum of the keys that land on it -- the
// smallest z wins and, among equal z, the smallest index, which is exactly
// the point the sequential loop keeps.  A resolve pass decodes z and gathers
// the winner's colour.  Three launches per call:
//   render_clear_kernel    z-buffer := all ones (empty)
//   render_splat_kernel    one lane per point: transform, project, atomicMin
//   render_resolve_kernel  one lane per pixel: depth, colour
#pragma clang fp contract(off)

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "recon_common.hpp"

namespace recon {

struct RenderParams {
    const float* xyz;       // 3 floats per point
    const uint8_t* valid;   // 1 byte per point, or null (all valid)
    const uint8_t* colors;  // 3 bytes per point, or null (white)
    int64_t num_points;
    float R[9];
    float t[3];
    float fx, fy, cx, cy;
    int32_t rows, cols;
};

constexpr unsigned long long kEmptyKey = ~0ull;

// Monotone map of float32 onto uint32: a < b  <=>  key(a) < key(b) for
// non-NaN values (-inf and +inf included; -0 sorts below +0, and z == 0 never
// reaches it).
__device__ __forceinline__ uint32_t depth_key(float z) {
    const uint32_t b = __float_as_uint(z);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_depth(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ __launch_bounds__(256) void render_clear_kernel(unsigned long long* __restrict__ zb,
                                                           int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) zb[i] = kEmptyKey;
}

// Camera-space z and z-buffer slot of one point; false when the point does
// not land on the image.
__device__ __forceinline__ bool splat_target(const RenderParams& p, float px, float py, float pz,
                                             float* z_out, int64_t* slot) {
    // world -> camera (se3.cuh:82-100)
    const float X = ((p.R[0] * px + p.R[1] * py) + p.R[2] * pz) + p.t[0];
    const float Y = ((p.R[3] * px + p.R[4] * py) + p.R[5] * pz) + p.t[1];
    const float Z = ((p.R[6] * px + p.R[7] * py) + p.R[8] * pz) + p.t[2];
    // z == 0 is skipped (cuda_point_cloud.cu:259-260); a z that is not below
    // FLT_MAX (FLT_MAX, +inf, NaN) never beats the FLT_MAX fill (:266-267)
    if (Z == 0.0f || !(Z < FLT_MAX)) return false;
    // static_cast<i32>((x * fx) / z + cx) truncates toward zero, so u in
    // (-1, cols) lands on columns 0..cols-1 (:262-265)
    const float u = (X * p.fx) / Z + p.cx;
    const float v = (Y * p.fy) / Z + p.cy;
    if (!(u > -1.0f && u < (float)p.cols && v > -1.0f && v < (float)p.rows)) return false;
    *z_out = Z;
    *slot = (int64_t)(int)v * p.cols + (int)u;
    return true;
}

// One point per lane.  The point and its mask byte are loaded together (the
// mask is a template switch, not a branch before the load).  The kernel is
// bound by the atomics: a pixel receives one per new minimum in arrival
// order (about ln k + 0.58 of k candidates for a random order), and wider
// lanes (several points each) measured slower, fewer waves hiding the
// atomics' latency.
template <bool kMask>
__global__ __launch_bounds__(256) void render_splat_kernel(const RenderParams p,
                                                           unsigned long long* __restrict__ zb) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.num_points) return;
    const float px = p.xyz[3 * i + 0], py = p.xyz[3 * i + 1], pz = p.xyz[3 * i + 2];
    const uint8_t ok = kMask ? p.valid[i] : uint8_t(1);
    float z;
    int64_t slot;
    if (!ok || !splat_target(p, px, py, pz, &z, &slot)) return;
    const unsigned long long key = ((unsigned long long)depth_key(z) << 32) | (uint32_t)i;
    // Keys only decrease during the pass, so a stale read is >= the current
    // minimum and skipping when it is <= key is safe; occluded points then
    // cost a load instead of an atomic.
    if (zb[slot] > key) atomicMin(zb + slot, key);
}

__global__ __launch_bounds__(256) void render_resolve_kernel(
    const RenderParams p, const unsigned long long* __restrict__ zb, float* __restrict__ depth,
    int32_t pitch, uint8_t* __restrict__ color, int64_t color_step) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= p.cols || y >= p.rows) return;
    const unsigned long long key = zb[(int64_t)y * p.cols + x];
    const bool hit = key != kEmptyKey;
    // filter_image_if(depth, FLT_MAX, 0) (cuda_point_cloud.cu:89-99)
    depth[(int64_t)y * pitch + x] = hit ? key_depth((uint32_t)(key >> 32)) : 0.0f;
    if (color) {
        uint8_t r = 0, g = 0, b = 0;
        if (hit) {
            if (p.colors) {
                const uint8_t* c = p.colors + 3 * (int64_t)(uint32_t)key;
                r = c[0];
                g = c[1];
                b = c[2];
            } else {
                r = g = b = 255;
            }
        }
        uint8_t* o = color + (int64_t)y * color_step + 3 * x;
        o[0] = r;
        o[1] = g;
        o[2] = b;
    }
}

}  // namespace recon

using namespace recon;

extern "C" int recon_points_render(const float* xyz, const uint8_t* valid, const uint8_t* colors,
                                   int64_t num_points, const double camera_pose[7], double fx,
                                   double fy, double cx, double cy, float* depth,
                                   size_t depth_step, int32_t rows, int32_t cols,
                                   uint8_t* color_out, size_t color_step,
                                   unsigned long long* zbuffer, recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_points_render";
    if (!camera_pose) return fail(RECON_ERR_NULL, "%s: camera_pose is NULL", fn);
    if (!depth) return fail(RECON_ERR_NULL, "%s: depth is NULL", fn);
    if (!zbuffer) return fail(RECON_ERR_NULL, "%s: zbuffer is NULL", fn);
    if (num_points < 0) return fail(RECON_ERR_SIZE, "%s: num_points %lld < 0", fn, (long long)num_points);
    if (num_points > 0 && !xyz) return fail(RECON_ERR_NULL, "%s: xyz is NULL", fn);
    if (num_points >= (int64_t(1) << 32))
        return fail(RECON_ERR_RANGE, "%s: %lld points exceed the 32-bit point index", fn,
                    (long long)num_points);
    if (rows <= 0 || cols <= 0)
        return fail(RECON_ERR_SIZE, "%s: image size %dx%d must be positive", fn, cols, rows);
    if (rows >= (1 << 18) || cols >= (1 << 24) || (int64_t)rows * cols >= (int64_t(1) << 31))
        return fail(RECON_ERR_RANGE, "%s: image size %dx%d too large", fn, cols, rows);
    if (depth_step < 4 * (size_t)cols || depth_step % 4 != 0)
        return fail(RECON_ERR_SIZE, "%s: depth step %zu invalid", fn, depth_step);
    if (depth_step / 4 >= (size_t(1) << 31))
        return fail(RECON_ERR_RANGE, "%s: depth step %zu too large", fn, depth_step);
    if (color_out && color_step < 3 * (size_t)cols)
        return fail(RECON_ERR_SIZE, "%s: color step %zu < 3*cols", fn, color_step);
    if (!std::isfinite((float)fx) || !std::isfinite((float)fy) || !std::isfinite((float)cx) ||
        !std::isfinite((float)cy) || (float)fx == 0.0f || (float)fy == 0.0f)
        return fail(RECON_ERR_CONFIG, "%s: invalid intrinsics", fn);
    if (!pose_is_finite(camera_pose)) return fail(RECON_ERR_CONFIG, "%s: non-finite pose", fn);
    int rc = require_gfx950();
    if (rc) return rc;
    RenderParams p;
    p.xyz = xyz;
    p.valid = valid;
    p.colors = colors;
    p.num_points = num_points;
    pose_to_matrix_f32(camera_pose, p.R, p.t);
    // the reference narrows the intrinsics to r32 (cuda_point_cloud.cu:248-251)
    p.fx = (float)fx;
    p.fy = (float)fy;
    p.cx = (float)cx;
    p.cy = (float)cy;
    p.rows = rows;
    p.cols = cols;
    const hipStream_t s = (hipStream_t)stream;
    const int64_t pixels = (int64_t)rows * cols;
    const int clear_blocks = (int)std::min<int64_t>((pixels + 255) / 256, 2048);
    hipLaunchKernelGGL(render_clear_kernel, dim3(clear_blocks), dim3(256), 0, s, zbuffer, pixels);
    RECON_HIP_TRY(hipGetLastError(), "render_clear_kernel launch");
    if (num_points > 0) {
        const dim3 grid((unsigned)((num_points + 255) / 256));
        if (valid)
            hipLaunchKernelGGL(render_splat_kernel<true>, grid, dim3(256), 0, s, p, zbuffer);
        else
            hipLaunchKernelGGL(render_splat_kernel<false>, grid, dim3(256), 0, s, p, zbuffer);
        RECON_HIP_TRY(hipGetLastError(), "render_splat_kernel launch");
    }
    hipLaunchKernelGGL(render_resolve_kernel, dim3((cols + 63) / 64, (rows + 3) / 4), dim3(256), 0,
                       s, p, (const unsigned long long*)zbuffer, depth, (int32_t)(depth_step / 4),
                       color_out, (int64_t)color_step);
    RECON_HIP_TRY(hipGetLastError(), "render_resolve_kernel launch");
    return RECON_OK;
}
