/*
 * points_oracle.c -- CPU restatement of the reference's depth -> point cloud
 * back-projection.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *   oracle_depth_to_points: point_cloud_extract<useRGB>
 *       (cu-Reconstruction modules/src/cuda/cuda_point_cloud.cu:22-70) with
 *       the float64 intrinsics it reads from the column-major K
 *       (cuda_point_cloud.cu:44-47, passed here as fx, fy, cx, cy).
 *   oracle_point_cloud_from_depth_f64: point_cloud_from_depth
 *       (modules/src/point_cloud_utility.cpp:76-115).
 */
#include "oracle.h"

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
