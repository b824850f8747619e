// mc_device.hpp -- device code of marching cubes shared by the single-volume
// path (marching_cubes.hip) and the batched path (tsdf_batch.hip):
// classification (marching_cubes.cu:76-84,118-130), edge intersection
// (marching_cubes.cu:23-51), the per-row triangle count and emission in the
// reference's output order, and a deterministic exclusive scan.
#pragma once

#include "recon_common.hpp"
#include "scan.hpp"

#include "recon/mc_tables.h"

namespace recon {

static __constant__ uint64_t c_tri[256] = RECON_MC_TRI_PACKED_INIT;
static __constant__ int c_edge[12][2] = RECON_MC_EDGE_CORNERS_INIT;

struct McParams {
    const float* f;
    int32_t gx, gy, gz;
    int32_t cx;        // cubes per row = gx - 1
    int32_t cyr;       // cube rows per layer = gy - 1
    int32_t rows;      // (gy - 1) * (gz - 1)
    float vs[3];
    float off[3];
};

// Corner i of cube (x,y,z) at voxel (x,y,z) + offset_i with offsets
// 0:(0,0,0) 1:(1,0,0) 2:(1,0,1) 3:(0,0,1) 4:(0,1,0) 5:(1,1,0) 6:(1,1,1)
// 7:(0,1,1) (marching_cubes.cu:118-127).
__device__ __forceinline__ void load_corners(const McParams& p, int x, int y, int z, float v[8]) {
    const int sy = p.gx, sz = p.gx * p.gy;
    const int i0 = x + sy * y + sz * z;
    v[0] = p.f[i0];
    v[1] = p.f[i0 + 1];
    v[2] = p.f[i0 + sz + 1];
    v[3] = p.f[i0 + sz];
    v[4] = p.f[i0 + sy];
    v[5] = p.f[i0 + sy + 1];
    v[6] = p.f[i0 + sy + sz + 1];
    v[7] = p.f[i0 + sy + sz];
}

// define_voxel_type (marching_cubes.cu:76-84)
__device__ __forceinline__ int cube_class(const float v[8]) {
    int c = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) c |= (v[i] < 0.0f ? 1 : 0) << i;
    return c;
}

__device__ __forceinline__ int class_ntri(int c) {
    return (c == 0 || c == 255) ? 0 : (int)(c_tri[c] & 0xFull);
}

// Sum over the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Inclusive prefix sum over the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_inclusive_scan(int v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Triangles of cube row r = y + (gy-1) * z (one wavefront; every lane returns
// the row total).
__device__ __forceinline__ int mc_row_triangles(const McParams& p, int r, int lane) {
    const int y = r % p.cyr, z = r / p.cyr;
    int n = 0;
    for (int x = lane; x < p.cx; x += 64) {
        float v[8];
        load_corners(p, x, y, z, v);
        n += class_ntri(cube_class(v));
    }
    return wave_sum(n);
}

// compute_intersection (marching_cubes.cu:23-51): orient the edge so that it
// runs from the negative corner when start > 0 and end < 0, then
// p = s + ((-ws) / (we - ws)) * (e - s), component by component.
__device__ __forceinline__ void edge_vertex(const McParams& p, int x, int y, int z,
                                            const float v[8], int e, float out[3]) {
    int a = c_edge[e][0], b = c_edge[e][1];
    float wa = v[a], wb = v[b];
    if ((wa > 0.0f) && (wb < 0.0f)) {
        const int ti = a; a = b; b = ti;
        const float tw = wa; wa = wb; wb = tw;
    }
    // corner offsets packed as bits: dx = bit0, dy = bit1, dz = bit2
    // (0:000 1:001 2:101 3:100 4:010 5:011 6:111 7:110)
    constexpr unsigned k_off = (0u << 0) | (1u << 3) | (5u << 6) | (4u << 9) | (2u << 12) |
                               (3u << 15) | (7u << 18) | (6u << 21);
    const unsigned oa = (k_off >> (3 * a)) & 7u, ob = (k_off >> (3 * b)) & 7u;
    const float ratio = (-wa) / (wb - wa);
    const int ia[3] = {x + (int)(oa & 1u), y + (int)((oa >> 1) & 1u), z + (int)(oa >> 2)};
    const int ib[3] = {x + (int)(ob & 1u), y + (int)((ob >> 1) & 1u), z + (int)(ob >> 2)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float s = voxel_centre(ia[k], p.vs[k], p.off[k]);
        const float en = voxel_centre(ib[k], p.vs[k], p.off[k]);
        const float edge = en - s;
        const float o = ratio * edge;
        out[k] = s + o;
    }
}

// Writes the triangles of cube row r starting at triangle `base` (one
// wavefront): a wave-level prefix sum of the per-cube counts places each
// cube's triangles in x order, slots in table order.  Triangles at or past
// `capacity` are not written.  Vertex indices (tris, may be null) are
// 3t+k + tri_index_base.
__device__ __forceinline__ void mc_emit_row(const McParams& p, int r, int lane, long long base,
                                            float* __restrict__ verts, int32_t* __restrict__ tris,
                                            long long capacity, long long tri_index_base) {
    const int y = r % p.cyr, z = r / p.cyr;
    for (int x0 = 0; x0 < p.cx; x0 += 64) {
        const int x = x0 + lane;
        float v[8];
        int c = 0, nt = 0;
        if (x < p.cx) {
            load_corners(p, x, y, z, v);
            c = cube_class(v);
            nt = class_ntri(c);
        }
        const int incl = wave_inclusive_scan(nt, lane);
        const int chunk_total = __shfl(incl, 63, 64);
        long long t = base + (incl - nt);
        if (nt) {
            const uint64_t word = c_tri[c];
            for (int s = 0; s < nt; ++s, ++t) {
                if (t >= capacity) break;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int e = (int)((word >> (4 + 4 * (3 * s + k))) & 0xFull);
                    float q[3];
                    edge_vertex(p, x, y, z, v, e, q);
                    float* dst = verts + 9 * t + 3 * k;
                    dst[0] = q[0];
                    dst[1] = q[1];
                    dst[2] = q[2];
                    if (tris) tris[3 * t + k] = (int32_t)(3 * (t - tri_index_base) + k);
                }
            }
        }
        base += chunk_total;
    }
}

}  // namespace recon
