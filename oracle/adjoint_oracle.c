/*
 * adjoint_oracle.c -- float64 CPU restatement of the differentiable integrate
 * (an extension of cu_tsdf_integrate, tsdf_volume.cu:144-190, with bilinear
 * depth sampling), of the surface loss built on the reference's trilinear
 * sampler (interpolate_trilinearly, tsdf_volume.cu:46-104) and of the adjoint
 * of both.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Poses are world->camera rotations R (row-major 3x3) and translations t in
 * float64 (12 doubles per frame); gradients are with respect to a left
 * perturbation T(xi) = Exp(xi) T, xi = (rho, omega): dp = rho + omega x p.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct sample {
    double p[3];      /* camera point */
    double u, v;      /* continuous pixel coordinates */
    int i0, j0;       /* top-left tap */
    double a, b;      /* fractional offsets */
    double d00, d10, d01, d11, d;
    int valid;        /* z > 0, 2x2 taps inside the image, all four valid */
} sample;

static inline void centre(const oracle_volume* vol, int64_t idx, double c[3]) {
    const int64_t gx = vol->grid[0], gy = vol->grid[1];
    const int64_t x = idx % gx, y = (idx / gx) % gy, z = idx / (gx * gy);
    c[0] = ((double)x + 0.5) * vol->voxel_size[0] + vol->offset[0];
    c[1] = ((double)y + 0.5) * vol->voxel_size[1] + vol->offset[1];
    c[2] = ((double)z + 0.5) * vol->voxel_size[2] + vol->offset[2];
}

static inline int tap_ok(const oracle_volume* vol, double d) {
    return d > 0.0 && d <= vol->max_depth;
}

/* Bilinear depth at the projection of world point c: depth[j][i] sits at
 * pixel coordinates (i, j), the ray of camera_util.cu:21-31. */
static void sample_depth(const oracle_volume* vol, const double* Rt, const float* depth,
                         int32_t rows, int32_t cols, const double c[3], sample* s) {
    for (int k = 0; k < 3; ++k)
        s->p[k] = Rt[3 * k] * c[0] + Rt[3 * k + 1] * c[1] + Rt[3 * k + 2] * c[2] + Rt[9 + k];
    s->valid = 0;
    if (!(s->p[2] > 0.0)) return;
    s->u = s->p[0] * vol->fx / s->p[2] + vol->cx;
    s->v = s->p[1] * vol->fy / s->p[2] + vol->cy;
    const double fu = floor(s->u), fv = floor(s->v);
    if (!(fu >= 0.0 && fu <= cols - 2 && fv >= 0.0 && fv <= rows - 2)) return;
    s->i0 = (int)fu;
    s->j0 = (int)fv;
    s->a = s->u - fu;
    s->b = s->v - fv;
    const float* r0 = depth + (int64_t)s->j0 * cols + s->i0;
    s->d00 = r0[0];
    s->d10 = r0[1];
    s->d01 = r0[cols];
    s->d11 = r0[cols + 1];
    if (!(tap_ok(vol, s->d00) && tap_ok(vol, s->d10) && tap_ok(vol, s->d01) && tap_ok(vol, s->d11)))
        return;
    s->d = (1 - s->a) * (1 - s->b) * s->d00 + s->a * (1 - s->b) * s->d10 +
           (1 - s->a) * s->b * s->d01 + s->a * s->b * s->d11;
    s->valid = 1;
}

void oracle_diff_forward_f64(const oracle_volume* vol, double* D, double* W, const float* depth,
                             int32_t rows, int32_t cols, const double* Rt, int32_t nframes,
                             int nthreads) {
    const int64_t n = (int64_t)vol->grid[0] * vol->grid[1] * vol->grid[2];
    const double trunc = vol->truncation, cap = vol->max_weight;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < n; ++i) {
        double c[3];
        centre(vol, i, c);
        for (int f = 0; f < nframes; ++f) {
            sample s;
            sample_depth(vol, Rt + 12 * f, depth + (int64_t)f * rows * cols, rows, cols, c, &s);
            if (!s.valid) continue;
            const double sdf = s.d - s.p[2];
            if (!(sdf >= -trunc)) continue;
            const double tsdf = sdf > 0 ? (sdf < trunc ? sdf : trunc) : sdf;
            const double w = W[i];
            D[i] = (D[i] * w + tsdf) / (w + 1.0);
            W[i] = (w + 1.0) < cap ? (w + 1.0) : cap;
        }
    }
}

/* The same forward in float32, one rounding per operation in the order
 * written (the GPU kernel's order): camera point ((r0 x + r1 y) + r2 z) + t,
 * s = ((p * f) / z) + c, footprint 0 <= s < size - 1, taps at (int)s,
 * d = (((ia ib) d00 + (a ib) d10) + (ia b) d01) + (a b) d11 with ia = 1 - a. */
void oracle_diff_forward_f32(const oracle_volume* vol, float* D, float* W, const float* depth,
                             int32_t rows, int32_t cols, const float* rt, int32_t nframes,
                             int nthreads) {
    const int64_t gx = vol->grid[0], gy = vol->grid[1], gz = vol->grid[2];
    const float trunc = vol->truncation, cap = vol->max_weight, maxd = vol->max_depth;
    const float fcols = (float)cols, frows = (float)rows;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
#endif
    for (int64_t z = 0; z < gz; ++z)
        for (int64_t y = 0; y < gy; ++y) {
            float c[3];
            for (int64_t x = 0; x < gx; ++x) {
                const int64_t i = x + gx * (y + gy * z);
                oracle_voxel_centre_f32(vol, (int32_t)x, (int32_t)y, (int32_t)z, c);
                for (int f = 0; f < nframes; ++f) {
                    const float* R = rt + 12 * f;
                    const float px = ((R[0] * c[0] + R[1] * c[1]) + R[2] * c[2]) + R[9];
                    const float py = ((R[3] * c[0] + R[4] * c[1]) + R[5] * c[2]) + R[10];
                    const float pz = ((R[6] * c[0] + R[7] * c[1]) + R[8] * c[2]) + R[11];
                    if (!(pz > 0.0f)) continue;
                    const float su = ((px * vol->fx) / pz) + vol->cx;
                    const float sv = ((py * vol->fy) / pz) + vol->cy;
                    if (!(su >= 0.0f && su < fcols - 1.0f && sv >= 0.0f && sv < frows - 1.0f)) continue;
                    const int i0 = (int)su, j0 = (int)sv;
                    const float a = su - (float)i0, b = sv - (float)j0;
                    const float* q = depth + (int64_t)f * rows * cols + (int64_t)j0 * cols + i0;
                    const float d00 = q[0], d10 = q[1], d01 = q[cols], d11 = q[cols + 1];
                    if (!(d00 > 0.0f && d00 <= maxd && d10 > 0.0f && d10 <= maxd && d01 > 0.0f &&
                          d01 <= maxd && d11 > 0.0f && d11 <= maxd))
                        continue;
                    const float ia = 1.0f - a, ib = 1.0f - b;
                    const float d = (((ia * ib) * d00 + (a * ib) * d10) + (ia * b) * d01) + (a * b) * d11;
                    const float sdf = d - pz;
                    if (!(sdf >= -trunc)) continue;
                    const float tsdf = sdf > 0.0f ? fminf(sdf, trunc) : sdf;
                    const float w = W[i], nw = w + 1.0f;
                    D[i] = ((D[i] * w) + tsdf) / nw;
                    W[i] = nw < cap ? nw : cap;
                }
            }
        }
}

/* Trilinear support of a world point (tsdf_volume.cu:46-104): grid-local q
 * (points more than 1e-4 below the grid move onto its face), lower corner
 * floor(q/vs - 0.5) clamped at 0, fractions (q - centre(lower)) / vs,
 * neighbour order 0:(0,0,0) 1:(1,0,0) 2:(1,0,1) 3:(0,0,1) 4:(0,1,0) 5:(1,1,0)
 * 6:(1,1,1) 7:(0,1,1).  The reference gives neighbour 6 the weight
 * (1-x)yz and neighbour 7 the weight xyz (tsdf_volume.cu:100-101), which
 * breaks interpolation of even an affine field; here each neighbour gets its
 * own trilinear weight.  Returns 0 when the 2x2x2 support leaves the grid. */
int oracle_trilinear_support(const oracle_volume* vol, const double pw[3], int64_t ids[8],
                             double coef[8]) {
    int lo[3];
    double t[3];
    for (int k = 0; k < 3; ++k) {
        /* grid-local coordinate with the reference's boundary clamps
         * (tsdf_volume.cu:53-64, BOUNDARY_EPS = 1e-4) */
        double q = pw[k] - vol->offset[k];
        const double bound = (double)vol->grid[k] * vol->voxel_size[k], eps = (double)1e-4f;
        if (q < -eps) q = 0.0;
        if (q > bound && q - bound < eps) q = bound - eps;
        if (!(q <= bound)) return 0; /* beyond the grid, or NaN */
        int l = (int)floor(q / vol->voxel_size[k] - 0.5);
        if (l < 0) l = 0;
        if (l + 1 > vol->grid[k] - 1) return 0;
        lo[k] = l;
        t[k] = (q - ((double)l + 0.5) * vol->voxel_size[k]) / vol->voxel_size[k];
    }
    const int64_t gx = vol->grid[0], gy = vol->grid[1];
    const int64_t dx = 1, dy = gx, dz = gx * gy;
    ids[0] = lo[0] + gx * (lo[1] + gy * (int64_t)lo[2]);
    ids[1] = ids[0] + dx;
    ids[2] = ids[1] + dz;
    ids[3] = ids[0] + dz;
    ids[4] = ids[0] + dy;
    ids[5] = ids[1] + dy;
    ids[6] = ids[2] + dy;
    ids[7] = ids[3] + dy;
    const double x = t[0], y = t[1], z = t[2];
    coef[0] = (1 - x) * (1 - y) * (1 - z);
    coef[1] = x * (1 - y) * (1 - z);
    coef[2] = x * (1 - y) * z;
    coef[3] = (1 - x) * (1 - y) * z;
    coef[4] = (1 - x) * y * (1 - z);
    coef[5] = x * y * (1 - z);
    coef[6] = x * y * z;       /* tsdf_volume.cu:100-101 swap these two */
    coef[7] = (1 - x) * y * z;
    return 1;
}

double oracle_surface_loss_f64(const oracle_volume* vol, const double* D, const double* points,
                               const double* weights, int64_t npoints, double* lambda) {
    const int64_t n = (int64_t)vol->grid[0] * vol->grid[1] * vol->grid[2];
    if (lambda) memset(lambda, 0, sizeof(double) * (size_t)n);
    double loss = 0.0;
    for (int64_t k = 0; k < npoints; ++k) {
        const double w = weights ? weights[k] : 1.0;
        if (w == 0.0) continue;
        int64_t ids[8];
        double coef[8];
        if (!oracle_trilinear_support(vol, points + 3 * k, ids, coef)) continue;
        double r = 0.0;
        for (int i = 0; i < 8; ++i) r += coef[i] * D[ids[i]];
        loss += w * r * r;
        if (lambda)
            for (int i = 0; i < 8; ++i) lambda[ids[i]] += 2.0 * w * r * coef[i] / (double)npoints;
    }
    return loss / (double)npoints;
}

void oracle_diff_adjoint_f64(const oracle_volume* vol, const float* depth, int32_t rows,
                             int32_t cols, const double* Rt, int32_t nframes, const double* lambda,
                             double* grad_pose, double* grad_depth) {
    const int64_t n = (int64_t)vol->grid[0] * vol->grid[1] * vol->grid[2];
    const double trunc = vol->truncation, cap = vol->max_weight;
    memset(grad_pose, 0, sizeof(double) * 6 * (size_t)nframes);
    if (grad_depth) memset(grad_depth, 0, sizeof(double) * (size_t)nframes * rows * cols);
    for (int64_t i = 0; i < n; ++i) {
        double lam = lambda[i];
        if (lam == 0.0) continue;
        double c[3];
        centre(vol, i, c);
        /* forward decisions of this voxel: count, then recompute in reverse */
        int cnt = 0;
        for (int f = 0; f < nframes; ++f) {
            sample s;
            sample_depth(vol, Rt + 12 * f, depth + (int64_t)f * rows * cols, rows, cols, c, &s);
            cnt += s.valid && (s.d - s.p[2]) >= -trunc;
        }
        /* reverse through the running average D' = a D + b tsdf */
        for (int f = nframes - 1; f >= 0 && cnt > 0; --f) {
            sample s;
            sample_depth(vol, Rt + 12 * f, depth + (int64_t)f * rows * cols, rows, cols, c, &s);
            if (!(s.valid && (s.d - s.p[2]) >= -trunc)) continue;
            cnt -= 1;
            const double wb = cnt < cap ? (double)cnt : cap;
            const double a = wb / (wb + 1.0), b = 1.0 / (wb + 1.0);
            const double sdf = s.d - s.p[2];
            const double g_sdf = lam * b * ((sdf > 0 && sdf >= trunc) ? 0.0 : 1.0);
            const double dd_du = (1 - s.b) * (s.d10 - s.d00) + s.b * (s.d11 - s.d01);
            const double dd_dv = (1 - s.a) * (s.d01 - s.d00) + s.a * (s.d11 - s.d10);
            const double z = s.p[2];
            const double g[3] = {g_sdf * dd_du * vol->fx / z, g_sdf * dd_dv * vol->fy / z,
                                 g_sdf * (-dd_du * vol->fx * s.p[0] / (z * z) -
                                          dd_dv * vol->fy * s.p[1] / (z * z) - 1.0)};
            double* gp = grad_pose + 6 * f;
            gp[0] += g[0];
            gp[1] += g[1];
            gp[2] += g[2];
            gp[3] += s.p[1] * g[2] - s.p[2] * g[1]; /* p x g */
            gp[4] += s.p[2] * g[0] - s.p[0] * g[2];
            gp[5] += s.p[0] * g[1] - s.p[1] * g[0];
            if (grad_depth) {
                double* gd = grad_depth + (int64_t)f * rows * cols + (int64_t)s.j0 * cols + s.i0;
                gd[0] += g_sdf * (1 - s.a) * (1 - s.b);
                gd[1] += g_sdf * s.a * (1 - s.b);
                gd[cols] += g_sdf * (1 - s.a) * s.b;
                gd[cols + 1] += g_sdf * s.a * s.b;
            }
            lam *= a;
        }
    }
}

/* 1 for points whose support voxels take every frame's decisions the same way
 * under any perturbation of size <= h of one pose coordinate (rho or omega of
 * one frame) or of one depth tap, and whose samples cross no integer pixel
 * line where the bilinear footprint moves with a slope jump above `kink`.
 * Per sample the camera point moves by at most e = 1.5 h (1 + |p|); the pixel
 * by du = fx e (1/z + |x|/z^2) / (1 - e/z) (likewise dv); the sdf by
 * 1.5 (Su du + Sv dv + e + h), Su and Sv the largest tap differences of the
 * window the footprint can reach. */
void oracle_stable_points(const oracle_volume* vol, const float* depth, int32_t rows,
                          int32_t cols, const double* Rt, int32_t nframes, const double* points,
                          int64_t npoints, double h, double kink, double* weights) {
    const double trunc = vol->truncation;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t k = 0; k < npoints; ++k) {
        int64_t ids[8];
        double coef[8];
        weights[k] = 0.0;
        if (!oracle_trilinear_support(vol, points + 3 * k, ids, coef)) continue;
        int ok = 1;
        for (int i = 0; i < 8 && ok; ++i) {
            double c[3];
            centre(vol, ids[i], c);
            for (int f = 0; f < nframes && ok; ++f) {
                const double* T = Rt + 12 * f;
                const float* dm = depth + (int64_t)f * rows * cols;
                double p[3];
                for (int q = 0; q < 3; ++q)
                    p[q] = T[3 * q] * c[0] + T[3 * q + 1] * c[1] + T[3 * q + 2] * c[2] + T[9 + q];
                const double e = 1.5 * h * (1.0 + sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
                if (p[2] < -e) continue;                      /* behind the camera */
                if (p[2] <= 4.0 * e) { ok = 0; break; }
                const double z = p[2], shrink = 1.0 / (1.0 - e / z);
                const double u = p[0] * vol->fx / z + vol->cx, v = p[1] * vol->fy / z + vol->cy;
                const double du = vol->fx * e * (1.0 / z + fabs(p[0]) / (z * z)) * shrink;
                const double dv = vol->fy * e * (1.0 / z + fabs(p[1]) / (z * z)) * shrink;
                /* columns / rows of every footprint the moved sample can use */
                const double xa = floor(u - du), xb = floor(u + du) + 1.0;
                const double ya = floor(v - dv), yb = floor(v + dv) + 1.0;
                if (xb - 1.0 < 0.0 || xa + 1.0 > cols - 1 || yb - 1.0 < 0.0 || ya + 1.0 > rows - 1)
                    continue;                                 /* no footprint inside */
                if (xa < 0.0 || xb > cols - 1 || ya < 0.0 || yb > rows - 1) { ok = 0; break; }
                const int x0 = (int)xa, x1 = (int)xb, y0 = (int)ya, y1 = (int)yb;
                int nvalid = 0, nmargin = 0, ntaps = 0;
                double su = 0.0, sv = 0.0;
                for (int y = y0; y <= y1; ++y)
                    for (int x = x0; x <= x1; ++x) {
                        const double d = dm[(int64_t)y * cols + x];
                        ++ntaps;
                        nvalid += tap_ok(vol, d);
                        nmargin += d > h && d <= vol->max_depth - h;
                        if (x < x1) su = fmax(su, fabs(dm[(int64_t)y * cols + x + 1] - d));
                        if (y < y1) sv = fmax(sv, fabs(dm[(int64_t)(y + 1) * cols + x] - d));
                    }
                if (nvalid == 0 && nmargin == 0) continue;    /* robustly no depth */
                if (nmargin < ntaps) { ok = 0; break; }
                /* slope jumps at the integer lines the sample crosses */
                for (int n = x0 + 1; n <= x1 - 1 && ok; ++n)
                    for (int y = y0; y <= y1; ++y) {
                        const float* r = dm + (int64_t)y * cols;
                        if (fabs((double)r[n + 1] - 2.0 * r[n] + r[n - 1]) > kink) { ok = 0; break; }
                    }
                for (int n = y0 + 1; n <= y1 - 1 && ok; ++n)
                    for (int x = x0; x <= x1; ++x) {
                        const double a = dm[(int64_t)(n - 1) * cols + x], b = dm[(int64_t)n * cols + x],
                                     cc = dm[(int64_t)(n + 1) * cols + x];
                        if (fabs(cc - 2.0 * b + a) > kink) { ok = 0; break; }
                    }
                if (!ok) break;
                sample s;
                sample_depth(vol, T, dm, rows, cols, c, &s);
                if (!s.valid) { ok = 0; break; }
                const double sdf = s.d - s.p[2];
                const double ds = 1.5 * (su * du + sv * dv + e + h);
                if (fabs(sdf + trunc) <= ds || fabs(sdf - trunc) <= ds) ok = 0;
            }
        }
        weights[k] = ok ? 1.0 : 0.0;
    }
}

/* Colour at points (float32, the device sampler's operation order): the
 * support of tri_support -- q = p - off with the BOUNDARY_EPS clamps,
 * vox = floor(q / vs), lower = vox - 1 when q lies below the voxel's centre,
 * at least 0, t = (q - centre(lower)) / vs -- and neighbours weighted by
 * coef * wc, c = sum(coef*wc*c) / sum(coef*wc) accumulated in neighbour
 * order, rounded to nearest (rintf) and clamped to 0..255. */
void oracle_sample_colors_f32(const oracle_volume* vol, const float* C, const float* points,
                              int64_t npoints, uint8_t* out) {
    const int gx = vol->grid[0], gy = vol->grid[1];
    for (int64_t k = 0; k < npoints; ++k) {
        const float* p = points + 3 * k;
        int lo[3];
        float t[3];
        int ok = 1;
        for (int a = 0; a < 3 && ok; ++a) {
            const float vs = vol->voxel_size[a];
            float q = p[a] - vol->offset[a];
            const float bound = (float)vol->grid[a] * vs;
            if (q < -1e-4f) q = 0.0f;
            if ((q > bound) && (q - bound < 1e-4f)) q = bound - 1e-4f;
            if (!(q <= bound)) { ok = 0; break; }
            const int vox = (int)floorf(q / vs);
            const float ct = ((float)vox + 0.5f) * vs;
            int l = q < ct ? vox - 1 : vox;
            if (l < 0) l = 0;
            if (!(l + 1 <= vol->grid[a] - 1)) { ok = 0; break; }
            lo[a] = l;
            const float lct = ((float)l + 0.5f) * vs;
            t[a] = (q - lct) / vs;
        }
        float r = 0.0f, g = 0.0f, b = 0.0f, wsum = 0.0f;
        if (ok) {
            const int64_t dx = 1, dy = gx, dz = (int64_t)gx * gy;
            int64_t ids[8];
            ids[0] = lo[0] + (int64_t)gx * (lo[1] + (int64_t)gy * lo[2]);
            ids[1] = ids[0] + dx;
            ids[2] = ids[1] + dz;
            ids[3] = ids[0] + dz;
            ids[4] = ids[0] + dy;
            ids[5] = ids[1] + dy;
            ids[6] = ids[2] + dy;
            ids[7] = ids[3] + dy;
            const float x = t[0], y = t[1], z = t[2];
            const float coef[8] = {(1 - x) * (1 - y) * (1 - z), x * (1 - y) * (1 - z),
                                   x * (1 - y) * z,             (1 - x) * (1 - y) * z,
                                   (1 - x) * y * (1 - z),       x * y * (1 - z),
                                   x * y * z,                   (1 - x) * y * z};
            for (int i = 0; i < 8; ++i) {
                const float* c = C + 4 * ids[i];
                const float a = coef[i] * c[3];
                r = r + a * c[0];
                g = g + a * c[1];
                b = b + a * c[2];
                wsum = wsum + a;
            }
        }
        uint8_t q[3] = {0, 0, 0};
        if (wsum > 0.0f) {
            const float v[3] = {r / wsum, g / wsum, b / wsum};
            for (int i = 0; i < 3; ++i) q[i] = (uint8_t)fminf(fmaxf(rintf(v[i]), 0.0f), 255.0f);
        }
        out[3 * k] = q[0];
        out[3 * k + 1] = q[1];
        out[3 * k + 2] = q[2];
    }
}

/* Model-to-frame loss of pose refinement (no counterpart in the reference):
 * each pair scores one frame's depth at the projection of one model point,
 * as a quadratic truncated at tau (tau <= 0: not truncated).
 * With g = 2 w r / npairs and the bilinear slopes of sample_depth,
 *   dr/dp = (dd/du fx/z, dd/dv fy/z, -(dd/du fx x + dd/dv fy y)/z^2 - 1),
 *   dL/drho += g dr/dp,  dL/domega += p x (g dr/dp)
 * (a camera point moves by rho + omega x p under the left perturbation). */
double oracle_frame_loss_f64(const oracle_volume* vol, const float* depth, int32_t rows,
                             int32_t cols, const double* Rt, int32_t nframes, const double* points,
                             const int32_t* pairs, const double* weights, int64_t npairs,
                             double tau, double* grad_pose, double* residuals) {
    double loss = 0.0;
    for (int64_t k = 0; k < npairs; ++k) {
        const int32_t f = pairs[2 * k], q = pairs[2 * k + 1];
        if (residuals) residuals[k] = 0.0;
        if (f < 0 || f >= nframes) continue;
        sample s;
        sample_depth(vol, Rt + 12 * f, depth + (int64_t)f * rows * cols, rows, cols,
                     points + 3 * (int64_t)q, &s);
        if (!s.valid) continue;
        const double r = s.d - s.p[2];
        const double w = weights ? weights[k] : 1.0;
        if (residuals) residuals[k] = r;
        if (tau > 0.0 && !(fabs(r) < tau)) {
            loss += w * tau * tau;
            continue;
        }
        loss += w * r * r;
        if (!grad_pose) continue;
        const double g = 2.0 * w * r / (double)npairs;
        const double dd_du = (1 - s.b) * (s.d10 - s.d00) + s.b * (s.d11 - s.d01);
        const double dd_dv = (1 - s.a) * (s.d01 - s.d00) + s.a * (s.d11 - s.d10);
        const double iz = 1.0 / s.p[2];
        const double gx = g * dd_du * vol->fx * iz;
        const double gy = g * dd_dv * vol->fy * iz;
        const double gz = -(gx * s.p[0] + gy * s.p[1]) * iz - g;
        double* o = grad_pose + 6 * f;
        o[0] += gx;
        o[1] += gy;
        o[2] += gz;
        o[3] += s.p[1] * gz - s.p[2] * gy;
        o[4] += s.p[2] * gx - s.p[0] * gz;
        o[5] += s.p[0] * gy - s.p[1] * gx;
    }
    return npairs > 0 ? loss / (double)npairs : 0.0;
}

/* Gauss-Newton system of the model-to-frame term (no counterpart in the
 * reference): per frame H = sum 2 w J J^T / npairs and g = sum 2 w r J /
 * npairs over the valid pairs with |r| < tau (every valid pair when
 * tau <= 0), J = (j, p x j) the residual's derivative with respect to xi.
 * With normals (world, one per point), j is the point-to-plane
 * linearisation of r = d(s) - z: the frame sees the plane through the point
 * with camera-frame normal n = R n_w, and moving the point by dp changes r by
 * -(z / (n . p)) n . dp, so j = -(z / (n . p)) n; pairs whose |n . p| is
 * below 0.1 |n| |p| (grazing views) add nothing.  Without normals j is dr/dp
 * of oracle_frame_loss_f64 (the bilinear depth slopes), and g is that
 * function's gradient.  H is (nframes, 6, 6), row-major, and g (nframes, 6)
 * (may be NULL), both added to. */
void oracle_frame_normal_matrix_f64(const oracle_volume* vol, const float* depth, int32_t rows,
                                    int32_t cols, const double* Rt, int32_t nframes,
                                    const double* points, const double* normals,
                                    const int32_t* pairs, const double* weights, int64_t npairs,
                                    double tau, double* H, double* grad) {
    for (int64_t k = 0; k < npairs; ++k) {
        const int32_t f = pairs[2 * k], q = pairs[2 * k + 1];
        if (f < 0 || f >= nframes) continue;
        sample s;
        const double* T = Rt + 12 * f;
        sample_depth(vol, T, depth + (int64_t)f * rows * cols, rows, cols,
                     points + 3 * (int64_t)q, &s);
        if (!s.valid) continue;
        const double r = s.d - s.p[2];
        if (tau > 0.0 && !(fabs(r) < tau)) continue;
        const double w = weights ? weights[k] : 1.0;
        double j[3];
        if (normals) {
            const double* nw = normals + 3 * (int64_t)q;
            double n[3];
            for (int i = 0; i < 3; ++i) n[i] = T[3 * i] * nw[0] + T[3 * i + 1] * nw[1] + T[3 * i + 2] * nw[2];
            const double np = n[0] * s.p[0] + n[1] * s.p[1] + n[2] * s.p[2];
            const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            const double pn = sqrt(s.p[0] * s.p[0] + s.p[1] * s.p[1] + s.p[2] * s.p[2]);
            if (!(fabs(np) >= 0.1 * nn * pn)) continue;
            const double c = -s.p[2] / np;
            for (int i = 0; i < 3; ++i) j[i] = c * n[i];
        } else {
            const double dd_du = (1 - s.b) * (s.d10 - s.d00) + s.b * (s.d11 - s.d01);
            const double dd_dv = (1 - s.a) * (s.d01 - s.d00) + s.a * (s.d11 - s.d10);
            const double iz = 1.0 / s.p[2];
            j[0] = dd_du * vol->fx * iz;
            j[1] = dd_dv * vol->fy * iz;
            j[2] = -(j[0] * s.p[0] + j[1] * s.p[1]) * iz - 1.0;
        }
        const double J[6] = {j[0], j[1], j[2], s.p[1] * j[2] - s.p[2] * j[1],
                             s.p[2] * j[0] - s.p[0] * j[2], s.p[0] * j[1] - s.p[1] * j[0]};
        const double c = 2.0 * w / (double)npairs;
        if (grad)
            for (int a = 0; a < 6; ++a) grad[6 * (int64_t)f + a] += c * r * J[a];
        double* o = H + 36 * (int64_t)f;
        for (int a = 0; a < 6; ++a)
            for (int b = a; b < 6; ++b) {
                const double e = c * J[a] * J[b];
                o[6 * a + b] += e;
                if (b != a) o[6 * b + a] += e;
            }
    }
}
