/*
 * points_oracle.c -- CPU restatement of the reference's depth -> point cloud
 * back-projection.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *   oracle_depth_to_points: point_cloud_extract<useRGB>
 *       (cu-Reconstruction modules/src/cuda/cuda_point_cloud.cu:22-70) with
 *       the float64 intrinsics it reads from the column-major K
 *       (cuda_point_cloud.cu:44-47, passed here as fx, fy, cx, cy).
 *   oracle_point_cloud_from_depth_f64: point_cloud_from_depth
 *       (modules/src/point_cloud_utility.cpp:76-115).
 *   oracle_render_points: the host z-buffer of
 *       cudaPointCloud::project_to_depth / project_to_color_depth
 *       (cuda_point_cloud.cu:229-353) followed by filter_image_if
 *       (cuda_point_cloud.cu:89-99), with explicit intrinsics in place of
 *       K[2], K[5] (cuda_point_cloud.cu:248-251, which read 0 from the
 *       column-major K the kernel above uses).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>

void oracle_depth_to_points(const float* depth, int32_t rows, int32_t cols, int64_t pitch,
                            double fx, double fy, double cx, double cy,
                            const double transform[7],
                            const uint8_t* rgb, int64_t rgb_pitch,
                            float* xyz, uint8_t* valid, uint8_t* colors_out) {
    float R[9], t[3];
    oracle_pose_matrix_f32(transform, R, t);
    for (int32_t y = 0; y < rows; ++y)
        for (int32_t x = 0; x < cols; ++x) {
            const int64_t i = (int64_t)y * cols + x;
            const float d = depth[(int64_t)y * pitch + x];
            if (d <= 0.0f) { /* cuda_point_cloud.cu:55-58 (left unwritten there) */
                xyz[3 * i] = xyz[3 * i + 1] = xyz[3 * i + 2] = 0.0f;
                valid[i] = 0;
                if (colors_out) colors_out[3 * i] = colors_out[3 * i + 1] = colors_out[3 * i + 2] = 0;
                continue;
            }
            /* cuda_point_cloud.cu:51-53: float(x) - cx promotes to double */
            const float px = (float)((((double)(float)x - cx) * (double)d) / fx);
            const float py = (float)((((double)(float)y - cy) * (double)d) / fy);
            const float pz = d;
            /* cuda_point_cloud.cu:60, se3.cuh:82-100 */
            xyz[3 * i + 0] = ((R[0] * px + R[1] * py) + R[2] * pz) + t[0];
            xyz[3 * i + 1] = ((R[3] * px + R[4] * py) + R[5] * pz) + t[1];
            xyz[3 * i + 2] = ((R[6] * px + R[7] * py) + R[8] * pz) + t[2];
            valid[i] = 1;
            if (colors_out) {
                for (int k = 0; k < 3; ++k)
                    colors_out[3 * i + k] = rgb ? rgb[(int64_t)y * rgb_pitch + 3 * x + k] : 255;
            }
        }
}

int64_t oracle_point_cloud_from_depth_f64(const float* depth, int32_t rows, int32_t cols,
                                          int64_t pitch, double fx, double fy, double cx,
                                          double cy, const double transform[7],
                                          double depth_min, double depth_max,
                                          double* xyz, int64_t* pixel_index,
                                          int64_t capacity) {
    double qw = transform[0], qx = transform[1], qy = transform[2], qz = transform[3];
    const double n = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
    qw /= n; qx /= n; qy /= n; qz /= n;
    const double R[9] = {
        1.0 - 2.0 * (qy * qy + qz * qz), 2.0 * (qx * qy - qw * qz),       2.0 * (qx * qz + qw * qy),
        2.0 * (qx * qy + qw * qz),       1.0 - 2.0 * (qx * qx + qz * qz), 2.0 * (qy * qz - qw * qx),
        2.0 * (qx * qz - qw * qy),       2.0 * (qy * qz + qw * qx),       1.0 - 2.0 * (qx * qx + qy * qy)};
    int64_t count = 0;
    for (int32_t v = 0; v < rows; ++v)
        for (int32_t u = 0; u < cols; ++u) {
            const float d = depth[(int64_t)v * pitch + u];
            if (d <= depth_min || d >= depth_max) continue; /* point_cloud_utility.cpp:94-95 */
            const double pz = (double)d;
            const double px = (u - cx) * pz / fx;
            const double py = (v - cy) * pz / fy;
            if (count < capacity) {
                xyz[3 * count + 0] = R[0] * px + R[1] * py + R[2] * pz + transform[4];
                xyz[3 * count + 1] = R[3] * px + R[4] * py + R[5] * pz + transform[5];
                xyz[3 * count + 2] = R[6] * px + R[7] * py + R[8] * pz + transform[6];
                if (pixel_index) pixel_index[count] = (int64_t)v * cols + u;
            }
            ++count;
        }
    return count;
}

void oracle_render_points(const float* xyz, const uint8_t* valid, const uint8_t* colors,
                          int64_t num_points, const double pose[7], float fx, float fy, float cx,
                          float cy, int32_t rows, int32_t cols, int64_t pitch, float* depth,
                          uint8_t* color_out, int64_t color_pitch) {
    float R[9], t[3];
    oracle_pose_matrix_f32(pose, R, t); /* SE3<r32>(quat, trans), :246 */
    /* cpu_depth.setTo(FLT_MAX); cpu_color.setTo(0) (:254-255, :318-319) */
    for (int32_t y = 0; y < rows; ++y)
        for (int32_t x = 0; x < cols; ++x) {
            depth[(int64_t)y * pitch + x] = FLT_MAX;
            if (color_out)
                for (int k = 0; k < 3; ++k) color_out[(int64_t)y * color_pitch + 3 * x + k] = 0;
        }
    for (int64_t i = 0; i < num_points; ++i) {
        if (valid && !valid[i]) continue; /* points without a depth sample */
        const float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
        /* transform * v->pos, world to camera (se3.cuh:82-100) */
        const float X = ((R[0] * px + R[1] * py) + R[2] * pz) + t[0];
        const float Y = ((R[3] * px + R[4] * py) + R[5] * pz) + t[1];
        const float Z = ((R[6] * px + R[7] * py) + R[8] * pz) + t[2];
        if (Z == 0.0f) continue; /* :259-260 */
        /* static_cast<i32>((x * fx) / z + cx) (:262-263): truncation toward
         * zero; a value outside int32 or NaN is treated as off the image
         * (x86's cvttss2si gives INT_MIN there, which the bounds test drops) */
        const float u = (X * fx) / Z + cx;
        const float v = (Y * fy) / Z + cy;
        if (!(u > -1.0f && u < 2147483648.0f) || !(v > -1.0f && v < 2147483648.0f)) continue;
        const int32_t ix = (int32_t)u, iy = (int32_t)v;
        if (ix < 0 || ix >= cols || iy < 0 || iy >= rows) continue; /* :264-265 */
        float* d = &depth[(int64_t)iy * pitch + ix];
        if (*d > Z) { /* strict: the first of equal depths stays (:266-269) */
            *d = Z;
            if (color_out)
                for (int k = 0; k < 3; ++k)
                    color_out[(int64_t)iy * color_pitch + 3 * ix + k] =
                        colors ? colors[3 * i + k] : 255;
        }
    }
    /* filter_image_if(depth, FLT_MAX, 0): |d - FLT_MAX| < EPSILON only holds
     * for FLT_MAX itself (EPSILON = 1e-10, util.cu:6-10; :89-99, :275-278) */
    for (int32_t y = 0; y < rows; ++y)
        for (int32_t x = 0; x < cols; ++x)
            if (depth[(int64_t)y * pitch + x] == FLT_MAX) depth[(int64_t)y * pitch + x] = 0.0f;
}
