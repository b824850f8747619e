/*
 * tsdf_oracle.c -- CPU restatement of the reference's TSDF volume init and
 * projective integrate (cu-Reconstruction modules/src/cuda/tsdf_volume.cu,
 * camera_util.cu, se3.cuh).  TEST INFRASTRUCTURE ONLY: loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * product library.  Build: -O3 -ffp-contract=off (no FMA contraction, IEEE
 * float32 operations one at a time, as written).
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* se3.cuh:42-57 (SE3(float4 quat, float3 trans)) and se3.cuh:124-140 (init),
 * with the double->float narrowing of tsdf_volume.cu:271-272 (quat is stored
 * as float4(qx,qy,qz,qw)). */
void oracle_pose_matrix_f32(const double pose[7], float R[9], float t[3]) {
    const float qw = (float)pose[0], qx = (float)pose[1], qy = (float)pose[2], qz = (float)pose[3];
    const float x = 2.0f * qx, y = 2.0f * qy, z = 2.0f * qz;
    const float wx = x * qw, wy = y * qw, wz = z * qw;
    const float xx = x * qx, xy = y * qx, xz = z * qx;
    const float yy = y * qy, yz = z * qy, zz = z * qz;
    R[0] = 1.0f - (yy + zz); R[1] = xy - wz;          R[2] = xz + wy;
    R[3] = xy + wz;          R[4] = 1.0f - (xx + zz); R[5] = yz - wx;
    R[6] = xz - wy;          R[7] = yz + wx;          R[8] = 1.0f - (xx + yy);
    t[0] = (float)pose[4]; t[1] = (float)pose[5]; t[2] = (float)pose[6];
}

/* tsdf_volume.cu:117-119: ((v + 0.5f) * voxel_size) + global_offset. */
void oracle_voxel_centre_f32(const oracle_volume* vol, int32_t x, int32_t y, int32_t z, float c[3]) {
    c[0] = (((float)x + 0.5f) * vol->voxel_size[0]) + vol->offset[0];
    c[1] = (((float)y + 0.5f) * vol->voxel_size[1]) + vol->offset[1];
    c[2] = (((float)z + 0.5f) * vol->voxel_size[2]) + vol->offset[2];
}

/* tsdf_volume.cu:227-254: W <- 0, D <- max_truncated_dist. */
void oracle_tsdf_init(const oracle_volume* vol, float* D, float* W) {
    const int64_t n = (int64_t)vol->grid[0] * vol->grid[1] * vol->grid[2];
    for (int64_t i = 0; i < n; ++i) {
        D[i] = vol->truncation;
        W[i] = 0.0f;
    }
}

/* Pixel coordinate of camera-space point p along one axis
 * (camera_util.cu:15-16): s = (p_axis * f) / p_z + c, index = (int)s.
 * The index is valid iff -1 < s < size, the same set as
 * 0 <= (int)s < size (tsdf_volume.cu:166) for every finite s; comparing in
 * float first keeps the cast defined for every input. */
static inline int pixel_index(float pa, float f, float pz, float c, int32_t size, int32_t* out) {
    const float s = ((pa * f) / pz) + c;
    if (!(s > -1.0f && s < (float)size)) return 0;
    *out = (int32_t)s;
    return 1;
}

static int64_t integrate_f32_impl(const oracle_volume* vol, float* D, float* W, float* C,
                                  const float* depth, int32_t rows, int32_t cols,
                                  int64_t pitch, const uint8_t* rgb, int64_t rgb_pitch,
                                  const double pose[7], uint8_t* update_mask, int nthreads);

int64_t oracle_tsdf_integrate_f32(const oracle_volume* vol, float* D, float* W,
                                  const float* depth, int32_t rows, int32_t cols,
                                  int64_t pitch, const double pose[7],
                                  uint8_t* update_mask, int nthreads) {
    return integrate_f32_impl(vol, D, W, NULL, depth, rows, cols, pitch, NULL, 0, pose, update_mask,
                              nthreads);
}

/* The colour integrate (no reference counterpart on the dense path, whose
 * integrate ignores colour; the octree's add_observation, octree.cu:118-137,
 * keeps a weighted colour average): C holds (r, g, b, wc) per voxel; a voxel
 * updated with -trunc <= sdf <= trunc also sets
 * c = ((c * wc) + rgb[v][u]) / (wc + 1) per channel, wc = min(wc + 1, maxw). */
int64_t oracle_tsdf_integrate_color_f32(const oracle_volume* vol, float* D, float* W, float* C,
                                        const float* depth, int32_t rows, int32_t cols,
                                        int64_t pitch, const uint8_t* rgb, int64_t rgb_pitch,
                                        const double pose[7], int nthreads) {
    return integrate_f32_impl(vol, D, W, C, depth, rows, cols, pitch, rgb, rgb_pitch, pose, NULL,
                              nthreads);
}

static int64_t integrate_f32_impl(const oracle_volume* vol, float* D, float* W, float* C,
                                  const float* depth, int32_t rows, int32_t cols,
                                  int64_t pitch, const uint8_t* rgb, int64_t rgb_pitch,
                                  const double pose[7], uint8_t* update_mask, int nthreads) {
    float R[9], t[3];
    oracle_pose_matrix_f32(pose, R, t);
    const int32_t gx = vol->grid[0], gy = vol->grid[1], gz = vol->grid[2];
    const int64_t nrows = (int64_t)gy * gz;
    const float trunc = vol->truncation, maxw = vol->max_weight, maxd = vol->max_depth;
    const float fx = vol->fx, fy = vol->fy, cx = vol->cx, cy = vol->cy;
    int64_t total = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : total)
#endif
    for (int64_t r = 0; r < nrows; ++r) {
        const int32_t vy = (int32_t)(r % gy), vz = (int32_t)(r / gy);
        int64_t idx = r * gx;
        for (int32_t vx = 0; vx < gx; ++vx, ++idx) {
            if (update_mask) update_mask[idx] = 0;
            float c[3];
            oracle_voxel_centre_f32(vol, vx, vy, vz, c);
            /* se3.cuh:82-100: rotate then translate */
            const float px = ((R[0] * c[0] + R[1] * c[1]) + R[2] * c[2]) + t[0];
            const float py = ((R[3] * c[0] + R[4] * c[1]) + R[5] * c[2]) + t[1];
            const float pz = ((R[6] * c[0] + R[7] * c[1]) + R[8] * c[2]) + t[2];
            if (!(pz > 0.0f)) continue;
            int32_t u, v;
            if (!pixel_index(px, fx, pz, cx, cols, &u)) continue;
            if (!pixel_index(py, fy, pz, cy, rows, &v)) continue;
            const float d = depth[(int64_t)v * pitch + u];
            if (d <= 0.0f || d > maxd) continue;          /* tsdf_volume.cu:169 */
            const float sdf = d - pz;                      /* camera-space z */
            if (!(sdf >= -trunc)) continue;                /* tsdf_volume.cu:174 */
            float tsdf = sdf;
            if (sdf > 0.0f) tsdf = fminf(sdf, trunc);      /* tsdf_volume.cu:176-177 */
            const float w = W[idx];
            const float dv = D[idx];
            float new_w, new_d;
            if (vol->weight_type == 1) {
                /* WEIGHT_BY_VARIANCE: the sample weighs sigma(1)^2 / sigma(d)^2 */
                const float e1 = 1.0f - vol->noise[2];
                const float s1 = vol->noise[0] + (vol->noise[1] * (e1 * e1));
                const float e = d - vol->noise[2];
                const float sg = vol->noise[0] + (vol->noise[1] * (e * e));
                const float wo = (s1 * s1) / (sg * sg);
                new_w = w + wo;
                new_d = ((dv * w) + (wo * tsdf)) / new_w;
            } else {
                new_w = w + 1.0f;
                new_d = ((dv * w) + tsdf) / new_w; /* tsdf_volume.cu:183 */
            }
            W[idx] = new_w < maxw ? new_w : maxw;
            D[idx] = new_d;
            if (C && rgb && sdf <= trunc) {
                float* c4 = C + 4 * idx;
                const uint8_t* px = rgb + (int64_t)v * rgb_pitch + 3 * (int64_t)u;
                const float wc = c4[3], nwc = wc + 1.0f;
                for (int k = 0; k < 3; ++k) c4[k] = ((c4[k] * wc) + (float)px[k]) / nwc;
                c4[3] = nwc < maxw ? nwc : maxw;
            }
            if (update_mask) update_mask[idx] = 1;
            total += 1;
        }
    }
    return total;
}

int64_t oracle_tsdf_integrate_f64(const oracle_volume* vol, double* D, double* W,
                                  const float* depth, int32_t rows, int32_t cols,
                                  int64_t pitch, const double pose[7],
                                  int64_t* decision_flips, int nthreads) {
    float R[9], t[3];
    oracle_pose_matrix_f32(pose, R, t);
    /* float64 rotation from the same float-narrowed quaternion */
    const double qw = (float)pose[0], qx = (float)pose[1], qy = (float)pose[2], qz = (float)pose[3];
    const double Rd[9] = {
        1.0 - 2.0 * (qy * qy + qz * qz), 2.0 * (qx * qy - qw * qz),       2.0 * (qx * qz + qw * qy),
        2.0 * (qx * qy + qw * qz),       1.0 - 2.0 * (qx * qx + qz * qz), 2.0 * (qy * qz - qw * qx),
        2.0 * (qx * qz - qw * qy),       2.0 * (qy * qz + qw * qx),       1.0 - 2.0 * (qx * qx + qy * qy)};
    const double td[3] = {t[0], t[1], t[2]};
    const int32_t gx = vol->grid[0], gy = vol->grid[1], gz = vol->grid[2];
    const int64_t nrows = (int64_t)gy * gz;
    const float trunc = vol->truncation, maxd = vol->max_depth;
    const double truncd = vol->truncation, maxwd = vol->max_weight;
    const float fx = vol->fx, fy = vol->fy, cx = vol->cx, cy = vol->cy;
    int64_t total = 0, flips = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : total, flips)
#endif
    for (int64_t r = 0; r < nrows; ++r) {
        const int32_t vy = (int32_t)(r % gy), vz = (int32_t)(r / gy);
        int64_t idx = r * gx;
        for (int32_t vx = 0; vx < gx; ++vx, ++idx) {
            /* float32 decisions */
            float c[3];
            oracle_voxel_centre_f32(vol, vx, vy, vz, c);
            const float px = ((R[0] * c[0] + R[1] * c[1]) + R[2] * c[2]) + t[0];
            const float py = ((R[3] * c[0] + R[4] * c[1]) + R[5] * c[2]) + t[1];
            const float pz = ((R[6] * c[0] + R[7] * c[1]) + R[8] * c[2]) + t[2];
            int32_t u = -1, v = -1;
            int pass = pz > 0.0f && pixel_index(px, fx, pz, cx, cols, &u) &&
                       pixel_index(py, fy, pz, cy, rows, &v);
            float d = 0.0f;
            if (pass) {
                d = depth[(int64_t)v * pitch + u];
                pass = !(d <= 0.0f || d > maxd) && ((d - pz) >= -trunc);
            }
            /* float64 values */
            const double cd[3] = {((double)vx + 0.5) * vol->voxel_size[0] + vol->offset[0],
                                  ((double)vy + 0.5) * vol->voxel_size[1] + vol->offset[1],
                                  ((double)vz + 0.5) * vol->voxel_size[2] + vol->offset[2]};
            const double pxd = Rd[0] * cd[0] + Rd[1] * cd[1] + Rd[2] * cd[2] + td[0];
            const double pyd = Rd[3] * cd[0] + Rd[4] * cd[1] + Rd[5] * cd[2] + td[1];
            const double pzd = Rd[6] * cd[0] + Rd[7] * cd[1] + Rd[8] * cd[2] + td[2];
            if (decision_flips) {
                int pass64 = 0;
                if (pzd > 0.0) {
                    const double su = pxd * fx / pzd + cx, sv = pyd * fy / pzd + cy;
                    if (su > -1.0 && su < cols && sv > -1.0 && sv < rows) {
                        const double d64 = depth[(int64_t)(int32_t)sv * pitch + (int32_t)su];
                        pass64 = !(d64 <= 0.0 || d64 > maxd) && ((d64 - pzd) >= -truncd);
                    }
                }
                flips += (pass64 != pass);
            }
            if (!pass) continue;
            const double sdf = (double)d - pzd;
            double tsdf = sdf;
            if (sdf > 0.0) tsdf = sdf < truncd ? sdf : truncd;
            const double w = W[idx];
            D[idx] = (D[idx] * w + tsdf) / (w + 1.0);
            W[idx] = (w + 1.0) < maxwd ? (w + 1.0) : maxwd;
            total += 1;
        }
    }
    if (decision_flips) *decision_flips = flips;
    return total;
}
