// scene.cpp -- synthetic depth frames (see scene.h).
#include "scene.h"

#include <algorithm>
#include <cmath>
#include <limits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in (0, 1]: 53 random bits.
inline double unit(uint64_t h) { return (double)((h >> 11) + 1) * (1.0 / 9007199254740992.0); }

void rotation(const double q_in[4], double R[9]) {
    double w = q_in[0], x = q_in[1], y = q_in[2], z = q_in[3];
    const double n = std::sqrt(w * w + x * x + y * y + z * z);
    w /= n; x /= n; y /= n; z /= n;
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
}

const double kInf = std::numeric_limits<double>::infinity();

// Smallest t > eps where o + t d enters the primitive (or crosses the plane).
double hit(const recon_scene_prim& pr, const double o[3], const double d[3]) {
    const double eps = 1e-9;
    switch (pr.kind) {
        case RECON_PRIM_SPHERE: {
            const double oc[3] = {o[0] - pr.p[0], o[1] - pr.p[1], o[2] - pr.p[2]};
            const double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
            const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - pr.p[3] * pr.p[3];
            const double disc = b * b - a * c;
            if (disc < 0) return kInf;
            const double s = std::sqrt(disc);
            double t = (-b - s) / a;
            if (t > eps) return t;
            t = (-b + s) / a;
            return t > eps ? t : kInf;
        }
        case RECON_PRIM_BOX: {
            double t0 = -kInf, t1 = kInf;
            for (int i = 0; i < 3; ++i) {
                if (d[i] == 0.0) {
                    if (o[i] < pr.p[i] || o[i] > pr.p[3 + i]) return kInf;
                    continue;
                }
                double a = (pr.p[i] - o[i]) / d[i], b = (pr.p[3 + i] - o[i]) / d[i];
                if (a > b) std::swap(a, b);
                t0 = std::max(t0, a);
                t1 = std::min(t1, b);
            }
            if (t0 > t1) return kInf;
            if (t0 > eps) return t0;
            return t1 > eps ? t1 : kInf;
        }
        case RECON_PRIM_PLANE_Z: {
            if (d[2] == 0.0) return kInf;
            const double t = (pr.p[0] - o[2]) / d[2];
            return t > eps ? t : kInf;
        }
        default:
            return kInf;
    }
}

double sdf_one(const recon_scene_prim& pr, const double p[3]) {
    switch (pr.kind) {
        case RECON_PRIM_SPHERE: {
            const double dx = p[0] - pr.p[0], dy = p[1] - pr.p[1], dz = p[2] - pr.p[2];
            return std::sqrt(dx * dx + dy * dy + dz * dz) - pr.p[3];
        }
        case RECON_PRIM_BOX: {
            double q[3], out2 = 0, inside = -kInf;
            for (int i = 0; i < 3; ++i) {
                const double c = 0.5 * (pr.p[i] + pr.p[3 + i]), h = 0.5 * (pr.p[3 + i] - pr.p[i]);
                q[i] = std::fabs(p[i] - c) - h;
                if (q[i] > 0) out2 += q[i] * q[i];
                inside = std::max(inside, q[i]);
            }
            return std::sqrt(out2) + std::min(inside, 0.0);
        }
        case RECON_PRIM_PLANE_Z:
            return p[2] - pr.p[0];
        default:
            return kInf;
    }
}

}  // namespace

extern "C" int recon_scene_render(const recon_scene_prim* prims, int32_t num_prims,
                                  const double* poses_cw, const uint64_t* frame_ids,
                                  int32_t num_frames, double fx, double fy, double cx, double cy,
                                  int32_t cols, int32_t rows, const recon_scene_noise* noise,
                                  float* out, int32_t nthreads) {
    if ((!prims && num_prims > 0) || !poses_cw || !out) return -1;
    if (num_frames < 0 || cols <= 0 || rows <= 0 || num_prims < 0) return -2;
    const int64_t npix = (int64_t)cols * rows;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    (void)nthreads;
#endif
    for (int32_t f = 0; f < num_frames; ++f) {
        const double* pose = poses_cw + 7 * f;
        double Rcw[9];
        rotation(pose, Rcw);
        const double* t = pose + 4;
        // camera centre C = -R^T t; world direction of camera ray = R^T dir
        double C[3];
        for (int i = 0; i < 3; ++i) C[i] = -(Rcw[0 * 3 + i] * t[0] + Rcw[1 * 3 + i] * t[1] + Rcw[2 * 3 + i] * t[2]);
        const uint64_t fid = frame_ids ? frame_ids[f] : (uint64_t)f;
        float* dst = out + (int64_t)f * npix;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
        for (int32_t v = 0; v < rows; ++v) {
            for (int32_t u = 0; u < cols; ++u) {
                const double dc[3] = {(u - cx) / fx, (v - cy) / fy, 1.0};
                double dw[3];
                for (int i = 0; i < 3; ++i)
                    dw[i] = Rcw[0 * 3 + i] * dc[0] + Rcw[1 * 3 + i] * dc[1] + Rcw[2 * 3 + i] * dc[2];
                double best = kInf;
                for (int32_t k = 0; k < num_prims; ++k) best = std::min(best, hit(prims[k], C, dw));
                // dc.z == 1, so the ray parameter is the camera-space z
                double z = std::isfinite(best) ? best : 0.0;
                if (noise && z > 0.0) {
                    const uint64_t key = splitmix64(noise->seed ^ splitmix64(fid * 0x100000001B3ull ^
                                                                             (uint64_t)(v * (int64_t)cols + u)));
                    const double u1 = unit(splitmix64(key ^ 1)), u2 = unit(splitmix64(key ^ 2));
                    const double u3 = unit(splitmix64(key ^ 3));
                    if (u3 <= noise->dropout) {
                        z = 0.0;
                    } else {
                        const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                        z += (noise->sigma0 + noise->sigma_z2 * z * z) * g;
                        if (z < 0.0) z = 0.0;
                    }
                }
                dst[(int64_t)v * cols + u] = (float)z;
            }
        }
    }
    return 0;
}

extern "C" double recon_scene_sdf(const recon_scene_prim* prims, int32_t num_prims, const double p[3]) {
    double best = kInf;
    for (int32_t k = 0; k < num_prims; ++k) best = std::min(best, sdf_one(prims[k], p));
    return best;
}
