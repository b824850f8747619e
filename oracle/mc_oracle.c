/*
 * mc_oracle.c -- CPU restatement of the reference's marching cubes
 * (cu-Reconstruction modules/src/cuda/marching_cubes.cu:23-236 with the
 * tables of modules/include/cuda/marching_cubes_tables.cuh).  TEST
 * INFRASTRUCTURE ONLY (see oracle.h).  The only change from the reference is
 * the layer stride gx*gy (marching_cubes.cu:211 uses gx+gy); corner positions
 * are the analytic voxel centres of tsdf_volume.cu:117-119.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

#include "recon/mc_tables.h"

#ifdef _OPENMP
#include <omp.h>
#endif

static const uint64_t k_tri[256] = RECON_MC_TRI_PACKED_INIT;
static const int k_edge[12][2] = RECON_MC_EDGE_CORNERS_INIT;
static const int k_corner[8][3] = RECON_MC_CORNER_OFFSET_INIT;

/* marching_cubes.cu:118-127: the 8 corner values of cube (x,y,z). */
static inline void corner_values(const float* f, int32_t gx, int32_t gy, int32_t x, int32_t y,
                                 int32_t z, float v[8]) {
    for (int i = 0; i < 8; ++i) {
        const int64_t idx = (int64_t)(x + k_corner[i][0]) +
                            (int64_t)gx * ((y + k_corner[i][1]) + (int64_t)gy * (z + k_corner[i][2]));
        v[i] = f[idx];
    }
}

/* marching_cubes.cu:76-84: bit i <- value[i] < 0. */
static inline int cube_class(const float v[8]) {
    int c = 0;
    for (int i = 0; i < 8; ++i)
        if (v[i] < 0.0f) c |= 1 << i;
    return c;
}

/* Triangles of a class; classes 0 and 255 are skipped (marching_cubes.cu:130). */
static inline int class_ntri(int c) {
    return (c == 0 || c == 255) ? 0 : RECON_MC_NTRI(k_tri[c]);
}

static int64_t layer_count(const float* f, const int32_t grid[3], int32_t z, int64_t* hist) {
    const int32_t gx = grid[0], gy = grid[1];
    int64_t n = 0;
    for (int32_t y = 0; y < gy - 1; ++y)
        for (int32_t x = 0; x < gx - 1; ++x) {
            float v[8];
            corner_values(f, gx, gy, x, y, z, v);
            const int c = cube_class(v);
            if (hist) hist[c] += 1;
            n += class_ntri(c);
        }
    return n;
}

int64_t oracle_mc_count(const float* field, const int32_t grid[3], int64_t* class_hist,
                        int nthreads) {
    const int32_t gz = grid[2];
    if (grid[0] < 2 || grid[1] < 2 || gz < 2) {
        if (class_hist)
            for (int i = 0; i < 256; ++i) class_hist[i] = 0;
        return 0;
    }
    int64_t total = 0;
    int64_t* hists = class_hist ? (int64_t*)calloc((size_t)(gz - 1) * 256, sizeof(int64_t)) : NULL;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads) reduction(+ : total)
#endif
    for (int32_t z = 0; z < gz - 1; ++z)
        total += layer_count(field, grid, z, hists ? hists + (int64_t)z * 256 : NULL);
    if (class_hist) {
        for (int i = 0; i < 256; ++i) class_hist[i] = 0;
        for (int32_t z = 0; z < gz - 1; ++z)
            for (int i = 0; i < 256; ++i) class_hist[i] += hists[(int64_t)z * 256 + i];
        free(hists);
    }
    return total;
}

/* marching_cubes.cu:23-51 compute_intersection, with corner positions from
 * the analytic voxel centres. */
static inline void edge_vertex(const float v[8], const float pos[8][3], int e, float out[3]) {
    int a = k_edge[e][0], b = k_edge[e][1];
    float sw = v[a], ew = v[b];
    const float* s = pos[a];
    const float* en = pos[b];
    if ((sw > 0.0f) && (ew < 0.0f)) {
        const float* tp = s; s = en; en = tp;
        const float tw = sw; sw = ew; ew = tw;
    }
    const float ratio = (-sw) / (ew - sw);
    for (int k = 0; k < 3; ++k) {
        const float edge = en[k] - s[k];
        const float off = ratio * edge;
        out[k] = s[k] + off;
    }
}

int64_t oracle_mc_extract(const float* field, const int32_t grid[3],
                          const float voxel_size[3], const float offset[3],
                          float* verts, int32_t* tris, int32_t* edge_ids,
                          int64_t* cube_ids, int64_t capacity, int nthreads) {
    const int32_t gx = grid[0], gy = grid[1], gz = grid[2];
    if (gx < 2 || gy < 2 || gz < 2) return 0;
    oracle_volume vol = {{gx, gy, gz}, {voxel_size[0], voxel_size[1], voxel_size[2]},
                         {offset[0], offset[1], offset[2]}, 0, 0, 0, 0, 0, 0, 0, 0, {0, 0, 0}};
    /* per-layer counts, then an exclusive scan gives each layer's first
     * triangle: the emission order is z (outer), y, x, slot
     * (marching_cubes.cu:163-188 inside the z loop of :213-225). */
    int64_t* base = (int64_t*)malloc(sizeof(int64_t) * (size_t)gz);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int32_t z = 0; z < gz - 1; ++z) base[z] = layer_count(field, grid, z, NULL);
    int64_t total = 0;
    for (int32_t z = 0; z < gz - 1; ++z) {
        const int64_t n = base[z];
        base[z] = total;
        total += n;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int32_t z = 0; z < gz - 1; ++z) {
        int64_t t = base[z];
        for (int32_t y = 0; y < gy - 1; ++y)
            for (int32_t x = 0; x < gx - 1; ++x) {
                float v[8];
                corner_values(field, gx, gy, x, y, z, v);
                const int c = cube_class(v);
                const int nt = class_ntri(c);
                if (nt == 0) continue;
                float pos[8][3];
                for (int i = 0; i < 8; ++i)
                    oracle_voxel_centre_f32(&vol, x + k_corner[i][0], y + k_corner[i][1],
                                            z + k_corner[i][2], pos[i]);
                const uint64_t word = k_tri[c];
                for (int s = 0; s < nt; ++s, ++t) {
                    if (t >= capacity) continue;
                    for (int k = 0; k < 3; ++k) {
                        const int e = RECON_MC_EDGE(word, 3 * s + k);
                        if (verts) edge_vertex(v, pos, e, verts + 9 * t + 3 * k);
                        if (tris) tris[3 * t + k] = (int32_t)(3 * t + k);
                        if (edge_ids) edge_ids[3 * t + k] = e;
                    }
                    if (cube_ids)
                        cube_ids[t] = (int64_t)x + (int64_t)(gx - 1) * (y + (int64_t)(gy - 1) * z);
                }
            }
    }
    free(base);
    return total;
}

/* ---- welded mesh ---------------------------------------------------------
 * No reference counterpart: the reference emits a triangle soup (its
 * per-triangle vert_remap, marching_cubes.cu:171-179, never shares a
 * vertex).  One vertex per grid edge whose end values differ in sign
 * (v < 0 against not), owned by its lower end voxel, numbered in the order
 * (voxel x + gx*(y + gy*z), axis x, y, z).  Triangles are the soup's, in the
 * soup's order, as indices into these vertices.  The vertex runs from the
 * negative end N to the other end M: r = (-dN) / (dM - dN), p = sN + r*(sM -
 * sN) per component -- the soup's own formula whenever it orients the edge
 * from N, i.e. except when dM is exactly +-0 and the cube lists M first.
 * The normal is the TSDF gradient, central differences divided by
 * (vs + vs) (one-sided, divided by vs, on the grid border), lerped from N to
 * M with the same r and normalised ((0,0,0) for a zero gradient). */

static inline float grad_axis(const float* f, const int32_t grid[3], int32_t x, int32_t y,
                              int32_t z, int a, const float vs[3]) {
    const int64_t sx = 1, sy = grid[0], sz = (int64_t)grid[0] * grid[1];
    const int64_t st = a == 0 ? sx : (a == 1 ? sy : sz);
    const int32_t c = a == 0 ? x : (a == 1 ? y : z);
    const int64_t i = (int64_t)x + sy * y + sz * z;
    if (c > 0 && c < grid[a] - 1) return (f[i + st] - f[i - st]) / (vs[a] + vs[a]);
    if (c == 0) return grid[a] > 1 ? (f[i + st] - f[i]) / vs[a] : 0.0f;
    return (f[i] - f[i - st]) / vs[a];
}

int64_t oracle_mc_weld(const float* field, const int32_t grid[3], const float voxel_size[3],
                       const float offset[3], float* verts, float* normals, int64_t vertex_capacity,
                       int32_t* tris, int64_t triangle_capacity, int64_t* num_vertices) {
    const int32_t gx = grid[0], gy = grid[1], gz = grid[2];
    *num_vertices = 0;
    if (gx < 2 || gy < 2 || gz < 2) return 0;
    oracle_volume vol = {{gx, gy, gz}, {voxel_size[0], voxel_size[1], voxel_size[2]},
                         {offset[0], offset[1], offset[2]}, 0, 0, 0, 0, 0, 0, 0, 0, {0, 0, 0}};
    const int64_t n = (int64_t)gx * gy * gz;
    int64_t* id = (int64_t*)malloc(sizeof(int64_t) * 3 * (size_t)n);
    int64_t nv = 0;
    for (int32_t z = 0; z < gz; ++z)
        for (int32_t y = 0; y < gy; ++y)
            for (int32_t x = 0; x < gx; ++x) {
                const int64_t i = (int64_t)x + (int64_t)gx * (y + (int64_t)gy * z);
                for (int a = 0; a < 3; ++a) {
                    id[3 * i + a] = -1;
                    const int32_t q[3] = {x + (a == 0), y + (a == 1), z + (a == 2)};
                    if (q[a] >= grid[a]) continue;
                    const int64_t j = (int64_t)q[0] + (int64_t)gx * (q[1] + (int64_t)gy * q[2]);
                    if ((field[i] < 0.0f) == (field[j] < 0.0f)) continue;
                    const int64_t v = nv++;
                    id[3 * i + a] = v;
                    if (v >= vertex_capacity) continue;
                    /* N = negative end */
                    const int neg_first = field[i] < 0.0f;
                    const int32_t pn[3] = {neg_first ? x : q[0], neg_first ? y : q[1], neg_first ? z : q[2]};
                    const int32_t pm[3] = {neg_first ? q[0] : x, neg_first ? q[1] : y, neg_first ? q[2] : z};
                    const float dn = neg_first ? field[i] : field[j];
                    const float dm = neg_first ? field[j] : field[i];
                    const float r = (-dn) / (dm - dn);
                    float sn[3], sm[3];
                    oracle_voxel_centre_f32(&vol, pn[0], pn[1], pn[2], sn);
                    oracle_voxel_centre_f32(&vol, pm[0], pm[1], pm[2], sm);
                    if (verts)
                        for (int k = 0; k < 3; ++k) {
                            const float e = sm[k] - sn[k];
                            const float o = r * e;
                            verts[3 * v + k] = sn[k] + o;
                        }
                    if (normals) {
                        float g[3];
                        for (int k = 0; k < 3; ++k) {
                            const float gn = grad_axis(field, grid, pn[0], pn[1], pn[2], k, voxel_size);
                            const float gm = grad_axis(field, grid, pm[0], pm[1], pm[2], k, voxel_size);
                            const float e = gm - gn;
                            const float o = r * e;
                            g[k] = gn + o;
                        }
                        const float len2 = (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2];
                        if (len2 > 0.0f) {
                            const float len = sqrtf(len2);
                            for (int k = 0; k < 3; ++k) normals[3 * v + k] = g[k] / len;
                        } else {
                            for (int k = 0; k < 3; ++k) normals[3 * v + k] = 0.0f;
                        }
                    }
                }
            }
    *num_vertices = nv;
    /* triangles in the soup's order (z, y, x, slot) */
    int64_t t = 0;
    for (int32_t z = 0; z < gz - 1; ++z)
        for (int32_t y = 0; y < gy - 1; ++y)
            for (int32_t x = 0; x < gx - 1; ++x) {
                float v[8];
                corner_values(field, gx, gy, x, y, z, v);
                const int c = cube_class(v);
                const int nt = class_ntri(c);
                const uint64_t word = k_tri[c];
                for (int s = 0; s < nt; ++s, ++t) {
                    if (t >= triangle_capacity || !tris) continue;
                    for (int k = 0; k < 3; ++k) {
                        const int e = RECON_MC_EDGE(word, 3 * s + k);
                        const int* ca = k_corner[k_edge[e][0]];
                        const int* cb = k_corner[k_edge[e][1]];
                        int a = 0;
                        while (ca[a] == cb[a]) ++a;
                        const int32_t ox = x + (ca[0] < cb[0] ? ca[0] : cb[0]);
                        const int32_t oy = y + (ca[1] < cb[1] ? ca[1] : cb[1]);
                        const int32_t oz = z + (ca[2] < cb[2] ? ca[2] : cb[2]);
                        const int64_t i = (int64_t)ox + (int64_t)gx * (oy + (int64_t)gy * oz);
                        tris[3 * t + k] = (int32_t)id[3 * i + a];
                    }
                }
            }
    free(id);
    return t;
}
