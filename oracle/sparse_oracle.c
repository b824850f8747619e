/*
 * sparse_oracle.c -- CPU restatement of the sparse (brick) volume.  TEST
 * INFRASTRUCTURE ONLY (see oracle.h).  No counterpart on the reference's
 * path (a dense grid, tsdf_volume.cu:194-254); the lattice and the per-voxel
 * update are the dense ones (tsdf_volume.cu:117-119,144-190 as restated in
 * tsdf_oracle.c), the bricks are 8x8x8 voxels, brick b = floor(v / 8).
 *
 * Allocation of one frame: every pixel (u, v) with 0 < d <= max_depth gives
 * the eight camera points of its corners (u or u+1, v or v+1) at depths
 * d - trunc and d + trunc, x = ((su - cx) * z) / fx, y = ((sv - cy) * z) / fy,
 * moved to the world with R^T (p - t) as ((R[k]*qx + R[3+k]*qy) + R[6+k]*qz);
 * their bounding box gives the voxel range floor((lo - off) / vs) - 1 ..
 * floor((hi - off) / vs) + 1 on each axis, and every brick of that range is
 * allocated (at most 8 per axis; within the allocation bounds when set).
 * Then every voxel of every allocated brick takes the dense update.  Bricks
 * live in a hash map of this file's own.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BRICK 8
#define BRICK_VOXELS 512

struct oracle_sparse {
    oracle_volume vol;
    int bounded;
    int32_t blo[3], bhi[3];  /* allocation bounds in bricks, inclusive */
    int64_t n, cap;          /* bricks, capacity of the arrays */
    int32_t* coords;         /* 3 per brick */
    float* D;                /* 512 per brick */
    float* W;
    int64_t* table;          /* brick index + 1, 0 when free */
    int64_t tsize;           /* power of two */
};

static uint64_t key_of(int32_t x, int32_t y, int32_t z) {
    return ((uint64_t)(uint32_t)x * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)y * 0xC2B2AE3D27D4EB4Full) ^
           ((uint64_t)(uint32_t)z * 0x165667B19E3779F9ull);
}

static int64_t find(const oracle_sparse* s, int32_t x, int32_t y, int32_t z) {
    int64_t h = (int64_t)(key_of(x, y, z) & (uint64_t)(s->tsize - 1));
    for (;;) {
        const int64_t e = s->table[h];
        if (e == 0) return -1;
        const int32_t* c = s->coords + 3 * (e - 1);
        if (c[0] == x && c[1] == y && c[2] == z) return e - 1;
        h = (h + 1) & (s->tsize - 1);
    }
}

static void rehash(oracle_sparse* s, int64_t tsize) {
    free(s->table);
    s->tsize = tsize;
    s->table = (int64_t*)calloc((size_t)tsize, sizeof(int64_t));
    for (int64_t i = 0; i < s->n; ++i) {
        const int32_t* c = s->coords + 3 * i;
        int64_t h = (int64_t)(key_of(c[0], c[1], c[2]) & (uint64_t)(tsize - 1));
        while (s->table[h]) h = (h + 1) & (tsize - 1);
        s->table[h] = i + 1;
    }
}

static void allocate(oracle_sparse* s, int32_t x, int32_t y, int32_t z) {
    if (find(s, x, y, z) >= 0) return;
    if (s->n == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 1024;
        s->coords = (int32_t*)realloc(s->coords, sizeof(int32_t) * 3 * (size_t)s->cap);
        s->D = (float*)realloc(s->D, sizeof(float) * BRICK_VOXELS * (size_t)s->cap);
        s->W = (float*)realloc(s->W, sizeof(float) * BRICK_VOXELS * (size_t)s->cap);
    }
    const int64_t i = s->n++;
    s->coords[3 * i] = x;
    s->coords[3 * i + 1] = y;
    s->coords[3 * i + 2] = z;
    for (int k = 0; k < BRICK_VOXELS; ++k) {
        s->D[i * BRICK_VOXELS + k] = s->vol.truncation;
        s->W[i * BRICK_VOXELS + k] = 0.0f;
    }
    if (2 * s->n > s->tsize) rehash(s, 2 * s->tsize);
    else {
        int64_t h = (int64_t)(key_of(x, y, z) & (uint64_t)(s->tsize - 1));
        while (s->table[h]) h = (h + 1) & (s->tsize - 1);
        s->table[h] = i + 1;
    }
}

static int32_t floor_div8(int32_t v) { return v >= 0 ? v / 8 : -((-v + 7) / 8); }

oracle_sparse* oracle_sparse_create(const oracle_volume* vol) {
    oracle_sparse* s = (oracle_sparse*)calloc(1, sizeof(oracle_sparse));
    s->vol = *vol;
    s->tsize = 4096;
    s->table = (int64_t*)calloc((size_t)s->tsize, sizeof(int64_t));
    return s;
}

void oracle_sparse_destroy(oracle_sparse* s) {
    if (!s) return;
    free(s->coords);
    free(s->D);
    free(s->W);
    free(s->table);
    free(s);
}

void oracle_sparse_set_bounds(oracle_sparse* s, const int32_t lo[3], const int32_t hi[3]) {
    s->bounded = lo && hi;
    if (!s->bounded) return;
    for (int k = 0; k < 3; ++k) {
        s->blo[k] = floor_div8(lo[k]);
        s->bhi[k] = floor_div8(hi[k] - 1);
    }
}

void oracle_sparse_allocate_box(oracle_sparse* s, const int32_t lo[3], const int32_t hi[3]) {
    for (int32_t z = floor_div8(lo[2]); z <= floor_div8(hi[2] - 1); ++z)
        for (int32_t y = floor_div8(lo[1]); y <= floor_div8(hi[1] - 1); ++y)
            for (int32_t x = floor_div8(lo[0]); x <= floor_div8(hi[0] - 1); ++x) allocate(s, x, y, z);
}

static inline int pixel_index(float pa, float f, float pz, float c, int32_t size, int32_t* out) {
    const float sc = ((pa * f) / pz) + c;
    if (!(sc > -1.0f && sc < (float)size)) return 0;
    *out = (int32_t)sc;
    return 1;
}

int64_t oracle_sparse_integrate(oracle_sparse* s, const float* depth, int32_t rows, int32_t cols,
                                int64_t pitch, const double pose[7]) {
    const oracle_volume* vol = &s->vol;
    float R[9], t[3];
    oracle_pose_matrix_f32(pose, R, t);
    const float trunc = vol->truncation, maxd = vol->max_depth, maxw = vol->max_weight;
    /* allocation */
    for (int32_t v = 0; v < rows; ++v)
        for (int32_t u = 0; u < cols; ++u) {
            const float d = depth[(int64_t)v * pitch + u];
            if (!(d > 0.0f && d <= maxd)) continue;
            const float z0 = d - trunc, z1 = d + trunc;
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int c = 0; c < 8; ++c) {
                const float su = (float)(u + (c & 1)), sv = (float)(v + ((c >> 1) & 1));
                const float z = (c & 4) ? z1 : z0;
                const float xc = ((su - vol->cx) * z) / vol->fx;
                const float yc = ((sv - vol->cy) * z) / vol->fy;
                const float qx = xc - t[0], qy = yc - t[1], qz = z - t[2];
                for (int k = 0; k < 3; ++k) {
                    const float w = ((R[k] * qx + R[3 + k] * qy) + R[6 + k] * qz);
                    lo[k] = fminf(lo[k], w);
                    hi[k] = fmaxf(hi[k], w);
                }
            }
            int32_t b0[3], b1[3];
            int ok = 1;
            for (int k = 0; k < 3; ++k) {
                const float a = floorf((lo[k] - vol->offset[k]) / vol->voxel_size[k]) - 1.0f;
                const float b = floorf((hi[k] - vol->offset[k]) / vol->voxel_size[k]) + 1.0f;
                if (!(a > -8388608.0f && b < 8388608.0f)) { ok = 0; break; }
                b0[k] = floor_div8((int32_t)a);
                b1[k] = floor_div8((int32_t)b);
                if (b1[k] - b0[k] >= 8) b1[k] = b0[k] + 7;
                if (s->bounded) {
                    if (b0[k] < s->blo[k]) b0[k] = s->blo[k];
                    if (b1[k] > s->bhi[k]) b1[k] = s->bhi[k];
                }
            }
            if (!ok) continue;
            for (int32_t bz = b0[2]; bz <= b1[2]; ++bz)
                for (int32_t by = b0[1]; by <= b1[1]; ++by)
                    for (int32_t bx = b0[0]; bx <= b1[0]; ++bx) allocate(s, bx, by, bz);
        }
    /* the dense update over every voxel of every brick */
    for (int64_t b = 0; b < s->n; ++b) {
        const int32_t* bc = s->coords + 3 * b;
        for (int k = 0; k < BRICK_VOXELS; ++k) {
            const int32_t x = bc[0] * BRICK + (k & 7), y = bc[1] * BRICK + ((k >> 3) & 7),
                          z = bc[2] * BRICK + (k >> 6);
            float c[3];
            oracle_voxel_centre_f32(vol, x, y, z, c);
            const float px = ((R[0] * c[0] + R[1] * c[1]) + R[2] * c[2]) + t[0];
            const float py = ((R[3] * c[0] + R[4] * c[1]) + R[5] * c[2]) + t[1];
            const float pz = ((R[6] * c[0] + R[7] * c[1]) + R[8] * c[2]) + t[2];
            if (!(pz > 0.0f)) continue;
            int32_t pu, pv;
            if (!pixel_index(px, vol->fx, pz, vol->cx, cols, &pu)) continue;
            if (!pixel_index(py, vol->fy, pz, vol->cy, rows, &pv)) continue;
            const float d = depth[(int64_t)pv * pitch + pu];
            if (d <= 0.0f || d > maxd) continue;
            const float sdf = d - pz;
            if (!(sdf >= -trunc)) continue;
            float tsdf = sdf;
            if (sdf > 0.0f) tsdf = fminf(sdf, trunc);
            float* D = s->D + b * BRICK_VOXELS + k;
            float* W = s->W + b * BRICK_VOXELS + k;
            const float w = *W, nw = w + 1.0f;
            *D = ((*D * w) + tsdf) / nw;
            *W = nw < maxw ? nw : maxw;
        }
    }
    return s->n;
}

int64_t oracle_sparse_num_bricks(const oracle_sparse* s) { return s->n; }

void oracle_sparse_export(const oracle_sparse* s, const int32_t origin[3], const int32_t size[3],
                          float* D, float* W) {
    for (int32_t k = 0; k < size[2]; ++k)
        for (int32_t j = 0; j < size[1]; ++j)
            for (int32_t i = 0; i < size[0]; ++i) {
                const int32_t x = origin[0] + i, y = origin[1] + j, z = origin[2] + k;
                const int64_t b = find(s, floor_div8(x), floor_div8(y), floor_div8(z));
                const int64_t o = i + (int64_t)size[0] * (j + (int64_t)size[1] * k);
                if (b < 0) {
                    if (D) D[o] = s->vol.truncation;
                    if (W) W[o] = 0.0f;
                } else {
                    const int64_t v = b * BRICK_VOXELS + ((z & 7) * 8 + (y & 7)) * 8 + (x & 7);
                    if (D) D[o] = s->D[v];
                    if (W) W[o] = s->W[v];
                }
            }
}

int64_t oracle_sparse_bricks(const oracle_sparse* s, int32_t* coords, float* D, float* W,
                             int64_t capacity) {
    const int64_t n = s->n < capacity ? s->n : capacity;
    if (coords) memcpy(coords, s->coords, sizeof(int32_t) * 3 * (size_t)n);
    if (D) memcpy(D, s->D, sizeof(float) * BRICK_VOXELS * (size_t)n);
    if (W) memcpy(W, s->W, sizeof(float) * BRICK_VOXELS * (size_t)n);
    return s->n;
}
