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

// The corners of the cubes of one cube row, 64 at a time, for a wavefront
// walking the row (every lane calls every step).  Each of the row's four
// voxel rows is read once, two steps ahead (the step at x0 issues the loads
// of x0+128 .. x0+191), so a step never waits on a load it has just issued.
// A cube's x+1 corners come from the next lane (a DPP wave shift) and lane
// 63's from lane 0 of the following 64 voxels.  The values are
// load_corners'.
struct CornerStream {
    const float* f;
    int gx;
    int rows[4];  // voxel rows (y,z), (y,z+1), (y+1,z), (y+1,z+1)
    float cur[4];
    float nx[4];

    __device__ __forceinline__ CornerStream(const McParams& p, int y, int z, int lane) {
        f = p.f;
        gx = p.gx;
        const int r0 = p.gx * y + p.gx * p.gy * z;
        rows[0] = r0;
        rows[1] = r0 + p.gx * p.gy;
        rows[2] = r0 + p.gx;
        rows[3] = r0 + p.gx + p.gx * p.gy;
#pragma unroll
        for (int k = 0; k < 4; ++k) cur[k] = lane < gx ? f[rows[k] + lane] : 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) nx[k] = lane + 64 < gx ? f[rows[k] + lane + 64] : 0.0f;
    }

    // Corners of cube (x0 + lane, y, z), x0 = 0, 64, 128, ... in turn;
    // defined for lanes with x0 + lane < gx - 1.
    __device__ __forceinline__ void step(int x0, int lane, float v[8]) {
        const int xn = x0 + 128 + lane;
        float nn[4], hi[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) nn[k] = xn < gx ? f[rows[k] + xn] : 0.0f;
        float nxt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) nxt[k] = nx[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            // wave_shl:1 -- lane i receives lane i+1
            const float sh = __int_as_float(
                __builtin_amdgcn_update_dpp(0, __float_as_int(cur[k]), 0x130, 0xF, 0xF, false));
            const float first = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nxt[k]), 0));
            hi[k] = lane == 63 ? first : sh;
        }
        v[0] = cur[0];
        v[1] = hi[0];
        v[2] = hi[1];
        v[3] = cur[1];
        v[4] = cur[2];
        v[5] = hi[2];
        v[6] = hi[3];
        v[7] = cur[3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cur[k] = nxt[k];
            nx[k] = nn[k];
        }
    }
};

// K consecutive cube rows (y .. y+K-1, z) of one wavefront: their 2(K+1)
// voxel rows, each read once, streamed like CornerStream (instead of 4 per
// cube row).
template <int K>
struct CornerStreamK {
    static constexpr int R = 2 * (K + 1);  // voxel rows: (y+i, z) and (y+i, z+1)
    const float* f;
    int gx;
    int rows[R];
    float cur[R];
    float nx[R];

    __device__ __forceinline__ CornerStreamK(const McParams& p, int y, int z, int lane) {
        f = p.f;
        gx = p.gx;
        const int sz = p.gx * p.gy;
        const int r0 = p.gx * y + sz * z;
#pragma unroll
        for (int i = 0; i <= K; ++i) {
            rows[2 * i] = r0 + i * p.gx;
            rows[2 * i + 1] = r0 + i * p.gx + sz;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) cur[k] = lane < gx ? f[rows[k] + lane] : 0.0f;
#pragma unroll
        for (int k = 0; k < R; ++k) nx[k] = lane + 64 < gx ? f[rows[k] + lane + 64] : 0.0f;
    }

    // v[i] = corners of cube (x0 + lane, y + i, z), i < K
    __device__ __forceinline__ void step(int x0, int lane, float v[K][8]) {
        const int xn = x0 + 128 + lane;
        float nn[R], hi[R];
#pragma unroll
        for (int k = 0; k < R; ++k) nn[k] = xn < gx ? f[rows[k] + xn] : 0.0f;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const float sh = __int_as_float(
                __builtin_amdgcn_update_dpp(0, __float_as_int(cur[k]), 0x130, 0xF, 0xF, false));
            const float first = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nx[k]), 0));
            hi[k] = lane == 63 ? first : sh;
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int a = 2 * i, b = 2 * i + 2;  // rows y+i and y+i+1 at z (z+1 at +1)
            v[i][0] = cur[a]; v[i][1] = hi[a]; v[i][2] = hi[a + 1]; v[i][3] = cur[a + 1];
            v[i][4] = cur[b]; v[i][5] = hi[b]; v[i][6] = hi[b + 1]; v[i][7] = cur[b + 1];
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            cur[k] = nx[k];
            nx[k] = nn[k];
        }
    }
};

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

// The lane owning element k of per-lane counts whose inclusive wave scan is
// `incl` (k < incl of lane 63): the number of lanes with incl <= k.  Every
// lane must call.
__device__ __forceinline__ int wave_owner(int incl, int k) {
    int pos = 0;
#pragma unroll
    for (int st = 32; st >= 1; st >>= 1)
        if (__shfl(incl, pos + st - 1, 64) <= k) pos += st;
    return min(pos, 63);
}

// Inclusive prefix sum over the 64 lanes of a wavefront, by DPP moves
// (no LDS round trips): row shifts build the prefix within each row of 16
// lanes, row broadcasts of lanes 15 and 31 carry it across rows.
__device__ __forceinline__ int wave_inclusive_scan(int v, int lane) {
    (void)lane;
    int t = v;
    t += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);   // row_shr:1
    t += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);   // row_shr:2
    t += __builtin_amdgcn_update_dpp(0, v, 0x113, 0xF, 0xF, true);   // row_shr:3
    t += __builtin_amdgcn_update_dpp(0, t, 0x114, 0xF, 0xE, true);   // row_shr:4, banks 1-3
    t += __builtin_amdgcn_update_dpp(0, t, 0x118, 0xF, 0xC, true);   // row_shr:8, banks 2-3
    t += __builtin_amdgcn_update_dpp(0, t, 0x142, 0xA, 0xF, false);  // row_bcast:15, rows 1, 3
    t += __builtin_amdgcn_update_dpp(0, t, 0x143, 0xC, 0xF, false);  // row_bcast:31, rows 2, 3
    return t;
}

// Bit of the 64-cube chunk starting at x0 in a row's chunk mask: chunk c for
// c < 31, and bit 31 for every chunk from 31 on.
__device__ __forceinline__ uint32_t chunk_bit(int x0) { return 1u << min(x0 >> 6, 31); }

// Bitwise OR over the 64 lanes of a wavefront.
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= (uint32_t)__shfl_xor((int)v, d, 64);
    return v;
}

// Triangles of cube row r = y + (gy-1) * z (one wavefront; every lane returns
// the row total).  With `mask`, also the row's chunk mask (chunk_bit of every
// chunk with a triangle).
__device__ __forceinline__ int mc_row_triangles(const McParams& p, int r, int lane,
                                                uint32_t* mask = nullptr) {
    const int y = r % p.cyr, z = r / p.cyr;
    int n = 0;
    uint32_t m = 0;
    CornerStream cs(p, y, z, lane);
    for (int x0 = 0; x0 < p.cx; x0 += 64) {
        float v[8];
        cs.step(x0, lane, v);
        const int t = x0 + lane < p.cx ? class_ntri(cube_class(v)) : 0;
        n += t;
        m |= t > 0 ? chunk_bit(x0) : 0u;
    }
    if (mask) *mask = wave_or(m);
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

// Writes the triangles of the 64 cubes (x0 + lane, y, z) with corners v
// starting at triangle `base` (one wavefront) and returns their number: a
// wave-level prefix sum of the per-cube counts places each cube's triangles
// in x order, slots in table order, and the triangles are spread over the 64
// lanes (not one lane per cube), so a cube with 5 triangles does not hold up
// the wave.  Triangles at or past `capacity` are not written.  Vertex indices
// (tris, may be null) are 3t+k + tri_index_base.
__device__ __forceinline__ int mc_emit_chunk(const McParams& p, int x0, int y, int z, int lane,
                                             const float v[8], long long base,
                                             float* __restrict__ verts,
                                             int32_t* __restrict__ tris, long long capacity,
                                             long long tri_index_base) {
    const int x = x0 + lane;
    int c = 0, nt = 0;
    if (x < p.cx) {
        c = cube_class(v);
        nt = class_ntri(c);
    }
    const int incl = wave_inclusive_scan(nt, lane);
    const int excl = incl - nt;
    const int chunk_total = __shfl(incl, 63, 64);
    // triangle k of the chunk (k = 0 .. chunk_total-1, in the soup's
    // order) is written by lane k % 64: its cube is the lane whose
    // triangles include k, its slot k minus that lane's first triangle
    for (int k0 = 0; k0 < chunk_total; k0 += 64) {
        const int k = k0 + lane;
        const int owner = wave_owner(incl, k);
        const int oc = __shfl(c, owner, 64);
        const int slot = k - __shfl(excl, owner, 64);
        float ov[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) ov[i] = __shfl(v[i], owner, 64);
        const long long t = base + k;
        if (k < chunk_total && t < capacity) {
            const uint64_t word = c_tri[oc];
            const int ox = x0 + owner;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int e = (int)((word >> (4 + 4 * (3 * slot + q))) & 0xFull);
                float pt[3];
                edge_vertex(p, ox, y, z, ov, e, pt);
                float* dst = verts + 9 * t + 3 * q;
                dst[0] = pt[0];
                dst[1] = pt[1];
                dst[2] = pt[2];
                if (tris) tris[3 * t + q] = (int32_t)(3 * (t - tri_index_base) + q);
            }
        }
    }
    return chunk_total;
}

// Writes the triangles of cube row r starting at triangle `base` (one
// wavefront), its 64-cube chunks in x order (mc_emit_chunk).  Only chunks
// that `mask` marks (chunk_bit; a superset of those holding a triangle) are
// read, their corners loaded directly (x and x+1 of each of the four voxel
// rows); the others hold no triangle and move nothing.
__device__ __forceinline__ void mc_emit_row_masked(const McParams& p, int r, int lane,
                                                   long long base, uint32_t mask,
                                                   float* __restrict__ verts,
                                                   int32_t* __restrict__ tris,
                                                   long long capacity) {
    const int y = r % p.cyr, z = r / p.cyr;
    const int sy = p.gx, sz = p.gx * p.gy;
    const int r0 = sy * y + sz * z;
    // voxel rows (y,z), (y,z+1), (y+1,z), (y+1,z+1), as CornerStream
    const int rows[4] = {r0, r0 + sz, r0 + sy, r0 + sy + sz};
    for (int x0 = 0; x0 < p.cx; x0 += 64) {
        if (!(mask & chunk_bit(x0))) continue;
        const int x = x0 + lane;
        float lo[4], hi[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lo[k] = x < p.gx ? p.f[rows[k] + x] : 0.0f;
            hi[k] = x + 1 < p.gx ? p.f[rows[k] + x + 1] : 0.0f;
        }
        const float v[8] = {lo[0], hi[0], hi[1], lo[1], lo[2], hi[2], hi[3], lo[3]};
        base += mc_emit_chunk(p, x0, y, z, lane, v, base, verts, tris, capacity, 0);
    }
}

}  // namespace recon
