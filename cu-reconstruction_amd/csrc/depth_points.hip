// depth_points.hip -- depth map -> world point cloud for gfx950.
//
// Replaces cudaPointCloud::extract_points / point_cloud_extract<useRGB>
// (cu-Reconstruction modules/src/cuda/cuda_point_cloud.cu:22-70,179-226).
// The reference keeps K in a __device__ double[3][3] filled by
// cudaMemcpyToSymbol per call and writes a packed 33-byte Vertex; here the
// intrinsics are kernel arguments and the output is three planes
// (xyz float32, a validity byte, optional rgb bytes), written by 64 x 4
// thread tiles so each wavefront stores one contiguous run of a row.
#pragma clang fp contract(off)

#include "recon_common.hpp"

namespace recon {

struct PointsParams {
    const float* depth;
    int32_t pitch;       // floats per depth row
    int32_t cols, rows;
    int32_t rgb_pitch;   // bytes per rgb row
    const uint8_t* rgb;
    double fx, fy, cx, cy;
    float R[9];
    float t[3];
};

__global__ __launch_bounds__(256) void depth_points_kernel(const PointsParams p,
                                                           float* __restrict__ xyz,
                                                           uint8_t* __restrict__ valid,
                                                           uint8_t* __restrict__ colors) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= p.cols || y >= p.rows) return;
    const int i = y * p.cols + x;
    const float d = p.depth[y * p.pitch + x];
    if (d <= 0.0f) {  // cuda_point_cloud.cu:55-58
        xyz[3 * i + 0] = 0.0f;
        xyz[3 * i + 1] = 0.0f;
        xyz[3 * i + 2] = 0.0f;
        valid[i] = 0;
        if (colors) colors[3 * i + 0] = colors[3 * i + 1] = colors[3 * i + 2] = 0;
        return;
    }
    // cuda_point_cloud.cu:51-53: (r32(x) - cx) promotes to double
    const float px = (float)((((double)(float)x - p.cx) * (double)d) / p.fx);
    const float py = (float)((((double)(float)y - p.cy) * (double)d) / p.fy);
    const float pz = d;
    // cuda_point_cloud.cu:60 (se3.cuh:82-100)
    xyz[3 * i + 0] = ((p.R[0] * px + p.R[1] * py) + p.R[2] * pz) + p.t[0];
    xyz[3 * i + 1] = ((p.R[3] * px + p.R[4] * py) + p.R[5] * pz) + p.t[1];
    xyz[3 * i + 2] = ((p.R[6] * px + p.R[7] * py) + p.R[8] * pz) + p.t[2];
    valid[i] = 1;
    if (colors) {
        if (p.rgb) {
            const uint8_t* src = p.rgb + (size_t)y * p.rgb_pitch + 3 * x;
            colors[3 * i + 0] = src[0];
            colors[3 * i + 1] = src[1];
            colors[3 * i + 2] = src[2];
        } else {
            colors[3 * i + 0] = colors[3 * i + 1] = colors[3 * i + 2] = 255;
        }
    }
}

}  // namespace recon

using namespace recon;

extern "C" int recon_depth_to_points(const recon_depth_view* depth, double fx, double fy, double cx,
                                     double cy, const double transform[7], const uint8_t* rgb,
                                     size_t rgb_step, float* xyz, uint8_t* valid,
                                     uint8_t* colors_out, recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_depth_to_points";
    if (!depth || !depth->data) return fail(RECON_ERR_NULL, "%s: depth is NULL", fn);
    if (!transform) return fail(RECON_ERR_NULL, "%s: transform is NULL", fn);
    if (!xyz || !valid) return fail(RECON_ERR_NULL, "%s: output is NULL", fn);
    if (depth->rows <= 0 || depth->cols <= 0)
        return fail(RECON_ERR_SIZE, "%s: depth size %dx%d must be positive", fn, depth->cols,
                    depth->rows);
    if (depth->step < 4 * (size_t)depth->cols || depth->step % 4 != 0)
        return fail(RECON_ERR_SIZE, "%s: depth step %zu invalid", fn, depth->step);
    if ((depth->step / 4) * (size_t)depth->rows >= (size_t(1) << 31) ||
        3 * (size_t)depth->cols * depth->rows >= (size_t(1) << 31))
        return fail(RECON_ERR_RANGE, "%s: image too large for 32-bit indexing", fn);
    if (rgb && (rgb_step < 3 * (size_t)depth->cols || rgb_step >= (size_t(1) << 31)))
        return fail(RECON_ERR_SIZE, "%s: rgb step %zu invalid", fn, rgb_step);
    if (!(fx > 0.0) || !(fy > 0.0) || !std::isfinite(fx) || !std::isfinite(fy) ||
        !std::isfinite(cx) || !std::isfinite(cy))
        return fail(RECON_ERR_CONFIG, "%s: invalid intrinsics", fn);
    if (!pose_is_finite(transform)) return fail(RECON_ERR_CONFIG, "%s: non-finite transform", fn);
    int rc = require_gfx950();
    if (rc) return rc;
    PointsParams p;
    p.depth = depth->data;
    p.pitch = (int32_t)(depth->step / 4);
    p.cols = depth->cols;
    p.rows = depth->rows;
    p.rgb = rgb;
    p.rgb_pitch = (int32_t)rgb_step;
    p.fx = fx;
    p.fy = fy;
    p.cx = cx;
    p.cy = cy;
    pose_to_matrix_f32(transform, p.R, p.t);
    const dim3 grid((p.cols + 63) / 64, (p.rows + 3) / 4);
    hipLaunchKernelGGL(depth_points_kernel, grid, dim3(256), 0, (hipStream_t)stream, p, xyz, valid,
                       colors_out);
    RECON_HIP_TRY(hipGetLastError(), "depth_points_kernel launch");
    return RECON_OK;
}
