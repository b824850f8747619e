// mc_device.hpp -- device code of marching cubes shared by the single-volume
// path (marching_cubes.hip) and the batched path (tsdf_batch.hip):
// classification (marching_cubes.cu:76-84,118-130), edge intersection
// (marching_cubes.cu:23-51), the chunk-count pass, the group records and the
// emission in the reference's output order.
#pragma once

#include "recon_common.hpp"
#include "scan.hpp"

#include "recon/mc_tables.h"

namespace recon {

static __constant__ uint64_t c_tri[256] = RECON_MC_TRI_PACKED_INIT;

// The edge table's corners packed 4 bits per edge into two immediates, so
// that a lane's edge lookup is a shift, not a load.
constexpr int kEdgeCorners[12][2] = RECON_MC_EDGE_CORNERS_INIT;
constexpr uint64_t pack_edge_corner(int w) {
    uint64_t m = 0;
    for (int e = 0; e < 12; ++e) m |= (uint64_t)kEdgeCorners[e][w] << (4 * e);
    return m;
}
constexpr uint64_t kEdgeA = pack_edge_corner(0), kEdgeB = pack_edge_corner(1);

// The triangle count of every class (the packed table's low 4 bits, at most
// 5) packed 3 bits per class, 21 classes per word, so that filling a
// workgroup's count table takes no memory access.
constexpr uint64_t kTriPacked[256] = RECON_MC_TRI_PACKED_INIT;
struct NtriWords {
    uint64_t w[13];
};
constexpr NtriWords pack_ntri() {
    NtriWords r{};
    for (int c = 1; c < 255; ++c) r.w[c / 21] |= (kTriPacked[c] & 0xFull) << (3 * (c % 21));
    return r;
}
constexpr NtriWords kNtri = pack_ntri();

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
// Inclusive max over the 64 lanes of a wavefront of non-negative values, by
// DPP moves (row shifts, then row broadcasts of lanes 15 and 31).
__device__ __forceinline__ int wave_inclusive_max(int v) {
    int t = v;
    t = max(t, __builtin_amdgcn_update_dpp(0, t, 0x111, 0xF, 0xF, true));  // row_shr:1
    t = max(t, __builtin_amdgcn_update_dpp(0, t, 0x112, 0xF, 0xF, true));  // row_shr:2
    t = max(t, __builtin_amdgcn_update_dpp(0, t, 0x114, 0xF, 0xF, true));  // row_shr:4
    t = max(t, __builtin_amdgcn_update_dpp(0, t, 0x118, 0xF, 0xF, true));  // row_shr:8
    t = max(t, __builtin_amdgcn_update_dpp(0, t, 0x142, 0xA, 0xF, false));  // row_bcast:15
    t = max(t, __builtin_amdgcn_update_dpp(0, t, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return t;
}

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

// compute_intersection (marching_cubes.cu:23-51): orient the edge so that it
// runs from the negative corner when start > 0 and end < 0, then
// p = s + ((-ws) / (we - ws)) * (e - s), component by component.
__device__ __forceinline__ void edge_vertex(const McParams& p, int x, int y, int z,
                                            const float v[8], int e, float out[3]) {
    int a = (int)((kEdgeA >> (4 * e)) & 0xFull), b = (int)((kEdgeB >> (4 * e)) & 0xFull);
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

// ---- chunk-count marching cubes (single-volume and batched paths) ----
// Tiling of a volume for the chunk-count pass (host and device).
struct McTiling {
    int32_t gpl;     // 4-row groups per layer = ceil((gy-1)/4)
    int32_t groups;  // gpl * (gz-1)
    int32_t tiles;   // ceil(groups / 4): one workgroup each
    int32_t nch;     // 64-cube chunks per cube row = ceil((gx-1)/64)
};

__host__ __device__ inline McTiling mc_tiling(int gx, int gy, int gz) {
    McTiling t{0, 0, 0, 0};
    if (gx < 2 || gy < 2 || gz < 2) return t;
    t.gpl = (gy - 1 + 3) / 4;
    t.groups = t.gpl * (gz - 1);
    t.tiles = (t.groups + 3) / 4;
    t.nch = (gx - 1 + 63) / 64;
    return t;
}

// ---- chunk count --------------------------------------------------------------
// A wavefront takes a group of kGroupRows consecutive cube rows of one layer
// (the layer's last group may be shorter); a workgroup (tile) takes
// kTileGroups consecutive groups, so a tile covers consecutive cube rows.
constexpr int kGroupRows = 4;  // cube rows per wavefront of the count pass
constexpr int kTileGroups = 4;  // groups (wavefronts) per tile
constexpr int kVoxelRows = 2 * (kGroupRows + 1);  // (y+i, z) and (y+i, z+1), i = 0..4

// Group q: its first cube row r0 = y + (gy-1)*z, and its number of rows k.
__device__ __forceinline__ void mc_group(const McParams& p, int gpl, int q, int& y, int& z,
                                         int& k) {
    z = q / gpl;
    y = kGroupRows * (q - z * gpl);
    k = min(kGroupRows, p.cyr - y);
}

// Class of the cube whose corner sign bits sit in byte b of the count
// pass's packed word: bit 2j of the word is voxel row j at x, bit 2j+1 at
// x+1, so a cube's byte holds corners 0, 1, 3, 2, 4, 5, 7, 6 in bit order.
__device__ __forceinline__ int qbyte_class(int b) {
    return (b & 0x33) | ((b & 0x04) << 1) | ((b & 0x08) >> 1) | ((b & 0x40) << 1) |
           ((b & 0x80) >> 1);
}

// Triangles of class c (class_ntri) from kNtri's immediates.
__device__ __forceinline__ int class_ntri_imm(int c) {
    const int wi = c / 21;
    uint64_t w = 0;
#pragma unroll
    for (int i = 0; i < 13; ++i) w = wi == i ? kNtri.w[i] : w;
    return (int)((w >> (3 * (c - 21 * wi))) & 7ull);
}

// Entry i of the count pass's LDS table: triangles of the class in byte i.
__device__ __forceinline__ uint8_t mc_ntab_qbyte(int i) {
    return (uint8_t)class_ntri_imm(qbyte_class(i));
}

// Sign bits (v < 0, so -0 and NaN give 0: define_voxel_type,
// marching_cubes.cu:76-84) of the kVoxelRows values, row j at bit 2j.
__device__ __forceinline__ uint32_t mc_sign_bits(const float v[kVoxelRows]) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < kVoxelRows; ++j) m |= (v[j] < 0.0f ? 1u : 0u) << (2 * j);
    return m;
}

// Triangles of each 64-cube chunk of the tile's cube rows into cc (chunk c
// of row r at r*nch + c), the tile's total into *tile_sum_out and, when
// tile_nne_out is not null, its number of chunks with triangles into
// *tile_nne_out.  The chunk's voxel
// rows are loaded two chunks ahead; a cube's x+1 corners are the next lane's
// sign bits (a DPP wave shift), lane 63's from the following chunk.  Loads
// past the row end read its last voxel, and rows past the layer's last
// voxel row read that row (both in bounds); their cubes are not counted.
// Tile b of a volume (a 256-thread workgroup; every thread calls).  ntab:
// the workgroup's LDS table of triangle counts by packed-word byte
// (mc_ntab_qbyte, filled before the call); wtot: 8 ints of LDS.
__device__ __forceinline__ void mc_count_tile(const McParams& p, const McTiling& tl, const int b,
                                              uint16_t* __restrict__ cc,
                                              const uint8_t* __restrict__ ntab, int* wtot,
                                              int32_t* __restrict__ tile_sum_out,
                                              int32_t* __restrict__ tile_nne_out) {
    const int gpl = tl.gpl, groups = tl.groups, nch = tl.nch;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int q = b * kTileGroups + wave;
    int total = 0, nne = 0;
    if (q < groups) {
        int y, z, k;
        mc_group(p, gpl, q, y, z, k);
        const int r0 = y + p.cyr * z;
        // uniform row bases; per-lane 32-bit byte offsets (loads take the
        // scalar-base form)
        const char* rowp[kVoxelRows];
#pragma unroll
        for (int i = 0; i <= kGroupRows; ++i) {
            const int yy = min(y + i, p.gy - 1);
            rowp[2 * i] = reinterpret_cast<const char*>(
                p.f + ((size_t)p.gx * ((size_t)yy + (size_t)p.gy * z)));
            rowp[2 * i + 1] = rowp[2 * i] + 4 * (size_t)p.gx * p.gy;
        }
        const uint32_t xlast = (uint32_t)(p.gx - 1);
        auto load = [&](int chunk, float v[kVoxelRows]) {
            const uint32_t off = min((uint32_t)(64 * chunk + lane), xlast) << 2;
#pragma unroll
            for (int j = 0; j < kVoxelRows; ++j)
                v[j] = *reinterpret_cast<const float*>(rowp[j] + off);
        };
        uint16_t* ccr = cc + (size_t)r0 * nch;
        // chunk c from its sign bits pc and the next chunk's pn
        auto count = [&](int c, uint32_t pc, uint32_t pn) {
            const uint32_t sh = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pc, 0x130, 0xF, 0xF,
                                                                      false);  // lane i <- i+1
            const uint32_t first = (uint32_t)__builtin_amdgcn_readlane((int)pn, 0);
            const uint32_t hbits = lane == 63 ? first : sh;
            uint32_t word = pc | (hbits << 1);
            if (64 * c + lane >= p.cx) word = 0;
            int n[kGroupRows];
#pragma unroll
            for (int i = 0; i < kGroupRows; ++i)
                n[i] = i < k ? (int)ntab[(word >> (4 * i)) & 0xFFu] : 0;
            const int any = n[0] | n[1] | n[2] | n[3];
            int cnt01 = 0, cnt23 = 0;
            if (__ballot(any != 0)) {
                const int s01 = wave_inclusive_scan(n[0] | (n[1] << 16), lane);
                const int s23 = wave_inclusive_scan(n[2] | (n[3] << 16), lane);
                cnt01 = __builtin_amdgcn_readlane(s01, 63);
                cnt23 = __builtin_amdgcn_readlane(s23, 63);
                total += n[0] + n[1] + n[2] + n[3];
            }
            if (lane < k) {
                const int w = lane < 2 ? cnt01 : cnt23;
                const int rowc = (lane & 1) ? (w >> 16) : (w & 0xFFFF);
                ccr[(size_t)lane * nch + c] = (uint16_t)rowc;
                nne += rowc > 0 ? 1 : 0;
            }
        };
        // two buffers, each chunk's loads issued two chunks ahead
        float va[kVoxelRows], vb[kVoxelRows];
        load(0, va);
        load(1, vb);
        uint32_t pc = mc_sign_bits(va);
        load(2, va);
        for (int c = 0; c < nch; c += 2) {
            uint32_t pn = mc_sign_bits(vb);  // chunk c+1
            load(c + 3, vb);
            count(c, pc, pn);
            if (c + 1 >= nch) break;
            pc = pn;
            pn = mc_sign_bits(va);  // chunk c+2
            load(c + 4, va);
            count(c + 1, pc, pn);
            pc = pn;
        }
    }
    total = wave_sum(total);
    nne = wave_sum(nne);
    if (lane == 0) {
        wtot[threadIdx.x >> 6] = total;
        wtot[4 + (threadIdx.x >> 6)] = nne;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *tile_sum_out = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        if (tile_nne_out) *tile_nne_out = wtot[4] + wtot[5] + wtot[6] + wtot[7];
    }
}


// Inclusive sums within each row of 16 lanes (DPP row shifts): lane 15 of a
// row ends with the row's total.
__device__ __forceinline__ int row16_inclusive_sum(int v) {
    int t = v;
    t += __builtin_amdgcn_update_dpp(0, t, 0x111, 0xF, 0xF, true);  // row_shr:1
    t += __builtin_amdgcn_update_dpp(0, t, 0x112, 0xF, 0xF, true);  // row_shr:2
    t += __builtin_amdgcn_update_dpp(0, t, 0x114, 0xF, 0xF, true);  // row_shr:4
    t += __builtin_amdgcn_update_dpp(0, t, 0x118, 0xF, 0xF, true);  // row_shr:8
    return t;
}

// Sign words (mc_sign_bits) of the four voxels x..x+3 a lane holds per row.
__device__ __forceinline__ void mc_sign_bits4(const float4 v[kVoxelRows], uint32_t m[4]) {
    m[0] = m[1] = m[2] = m[3] = 0;
#pragma unroll
    for (int j = 0; j < kVoxelRows; ++j) {
        m[0] |= (v[j].x < 0.0f ? 1u : 0u) << (2 * j);
        m[1] |= (v[j].y < 0.0f ? 1u : 0u) << (2 * j);
        m[2] |= (v[j].z < 0.0f ? 1u : 0u) << (2 * j);
        m[3] |= (v[j].w < 0.0f ? 1u : 0u) << (2 * j);
    }
}

// mc_count_tile with 16-byte loads, for rows whose every 4-voxel run is
// 16-byte aligned (gx % 4 == 0 and a 16-byte-aligned field): the same
// outputs.  Lane l takes cubes x = 256 s + 4 l + q (q < 4) of span s, so
// chunk 4 s + m is the 16 lanes of DPP row m and each load instruction
// moves 1 KB of one voxel row.  Spans go in pairs, both loaded at once (one
// memory round trip per pair).  A cube's x+1 corners are the lane's next
// word, or for q = 3 the next lane's first (a DPP wave shift); lane 63's
// come from the pair's second span, or for the second span from voxels at
// x = 256 (s + 1) that lane 63 loads itself when that span exists.  In a
// row's last span lane 63's last cube (x = 256 s + 255) is past the row's
// last cube (64 nch >= gx - 1, and gx % 4 == 0 rules out gx - 1 = 256 nsp),
// so it needs no next span.  Loads past the row end read its last 16 bytes
// (in bounds; their cubes are not counted).
__device__ __forceinline__ void mc_count_tile4(const McParams& p, const McTiling& tl, const int b,
                                               uint16_t* __restrict__ cc,
                                               const uint8_t* __restrict__ ntab, int* wtot,
                                               int32_t* __restrict__ tile_sum_out,
                                               int32_t* __restrict__ tile_nne_out) {
    const int gpl = tl.gpl, groups = tl.groups, nch = tl.nch;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int q = b * kTileGroups + wave;
    int total = 0, nne = 0;
    if (q < groups) {
        int y, z, k;
        mc_group(p, gpl, q, y, z, k);
        const int r0 = y + p.cyr * z;
        const char* rowp[kVoxelRows];
#pragma unroll
        for (int i = 0; i <= kGroupRows; ++i) {
            const int yy = min(y + i, p.gy - 1);
            rowp[2 * i] = reinterpret_cast<const char*>(
                p.f + ((size_t)p.gx * ((size_t)yy + (size_t)p.gy * z)));
            rowp[2 * i + 1] = rowp[2 * i] + 4 * (size_t)p.gx * p.gy;
        }
        const uint32_t xlast4 = (uint32_t)(p.gx - 4), xlast = (uint32_t)(p.gx - 1);
        const int nsp = (nch + 3) / 4;
        uint16_t* ccr = cc + (size_t)r0 * nch;
        auto load = [&](int sp, float4 v[kVoxelRows]) {
            const uint32_t off = min((uint32_t)(256 * sp + 4 * lane), xlast4) << 2;
#pragma unroll
            for (int j = 0; j < kVoxelRows; ++j)
                v[j] = *reinterpret_cast<const float4*>(rowp[j] + off);
        };
        // span sp from its words pc and lane 63's next word h63
        auto count = [&](int sp, const uint32_t pc[4], uint32_t h63) {
            const uint32_t sh = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pc[0], 0x130, 0xF, 0xF,
                                                                      false);  // lane i <- i+1
            const uint32_t h3 = lane == 63 ? h63 : sh;
            const int xb = 256 * sp + 4 * lane;
            int n01 = 0, n23 = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t hb = e < 3 ? pc[e < 3 ? e + 1 : 0] : h3;
                uint32_t word = pc[e] | (hb << 1);
                if (xb + e >= p.cx) word = 0;
                int n[kGroupRows];
#pragma unroll
                for (int i = 0; i < kGroupRows; ++i)
                    n[i] = i < k ? (int)ntab[(word >> (4 * i)) & 0xFFu] : 0;
                n01 += n[0] | (n[1] << 16);
                n23 += n[2] | (n[3] << 16);
            }
            total += (n01 & 0xFFFF) + (n01 >> 16) + (n23 & 0xFFFF) + (n23 >> 16);
            // per chunk and row at most 64 cubes x 5 triangles: 16-bit fields
            n01 = row16_inclusive_sum(n01);
            n23 = row16_inclusive_sum(n23);
            const int c = 4 * sp + (lane >> 4);
            if ((lane & 15) == 15 && c < nch) {
#pragma unroll
                for (int i = 0; i < kGroupRows; ++i) {
                    if (i >= k) break;
                    const int w = i < 2 ? n01 : n23;
                    const int rowc = (i & 1) ? (w >> 16) : (w & 0xFFFF);
                    ccr[(size_t)i * nch + c] = (uint16_t)rowc;
                    nne += rowc > 0 ? 1 : 0;
                }
            }
        };
        for (int sp = 0; sp < nsp; sp += 2) {
            // a lone last span loads its own voxels twice (clamped, in bounds)
            float4 va[kVoxelRows], vb[kVoxelRows];
            load(sp, va);
            load(sp + 1 < nsp ? sp + 1 : sp, vb);
            uint32_t hn = 0;
            if (lane == 63 && sp + 2 < nsp) {
                const uint32_t offn = min((uint32_t)(256 * sp + 512), xlast) << 2;
#pragma unroll
                for (int j = 0; j < kVoxelRows; ++j)
                    hn |= (*reinterpret_cast<const float*>(rowp[j] + offn) < 0.0f ? 1u : 0u)
                          << (2 * j);
            }
            uint32_t pa[4], pb[4];
            mc_sign_bits4(va, pa);
            mc_sign_bits4(vb, pb);
            count(sp, pa, (uint32_t)__builtin_amdgcn_readlane((int)pb[0], 0));
            if (sp + 1 < nsp) count(sp + 1, pb, hn);
        }
    }
    total = wave_sum(total);
    nne = wave_sum(nne);
    if (lane == 0) {
        wtot[threadIdx.x >> 6] = total;
        wtot[4 + (threadIdx.x >> 6)] = nne;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        *tile_sum_out = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        if (tile_nne_out) *tile_nne_out = wtot[4] + wtot[5] + wtot[6] + wtot[7];
    }
}
// ---- emit -----------------------------------------------------------------------
// Per-workgroup LDS of the emit pass.
struct McEmitLds {
    uint64_t tri[256];              // c_tri
    int lchunk[256], loff[256];     // compacted chunks with triangles, their first triangle
    int wtot[4];
    uint8_t ntab[256];              // triangles by class
    float stage[4][64 * 9];         // per wavefront: 64 triangles' vertices
    int gchunk[4][4], goff[4][4];   // per wavefront: its group's chunks, first triangles
    int mark[4][64];                // per wavefront: the triangle owners' marks
};

// Fills the tables of an emit workgroup (256 threads); the first barrier of
// mc_emit_tile orders them before their use.
__device__ __forceinline__ void mc_emit_lds_init(McEmitLds& l) {
    l.tri[threadIdx.x] = c_tri[threadIdx.x];
    l.ntab[threadIdx.x] = (uint8_t)class_ntri((int)threadIdx.x);
}

// Writes the triangles of up to four chunks (list entries e0 .. e0+ne-1 of
// the tile's compacted list, consecutive in the output from triangle t0) by
// one wavefront: lane l takes cubes 4(l%16) .. +3 of chunk l/16, so the
// lanes' cubes are in output order; a wave-level prefix sum places each
// lane's triangles, and triangle k of the group is computed by lane k % 64,
// which finds its cube (wave_owner, then among the owner's four), takes the
// owner's 20 voxel values (wave shuffles; the cube's 8 corners are among
// them) and computes the vertices (edge_vertex) into an LDS stage; each 64
// triangles' vertices and indices then leave in consecutive dwords.
__device__ __forceinline__ void mc_emit_group(const McParams& p, int nch, const int* lchunk,
                                              int ne, long long t0, int lane,
                                              const McEmitLds& l, float* lstage, int* mark,
                                              float* __restrict__ verts,
                                              int32_t* __restrict__ tris, long long capacity) {
    const int slot4 = lane >> 4;
    const bool act = slot4 < ne;
    const int chunk = lchunk[act ? slot4 : 0];
    const int r = chunk / nch;
    const int x = 64 * (chunk - r * nch) + 4 * (lane & 15);
    const int y = r % p.cyr, z = r / p.cyr;
    const size_t sy = (size_t)p.gx, sz = (size_t)p.gx * p.gy;
    const size_t i0 = (size_t)y * sy + (size_t)z * sz;
    const float* rp[4] = {p.f + i0, p.f + i0 + sz, p.f + i0 + sy, p.f + i0 + sy + sz};
    float v[4][5];
    const int xlast = p.gx - 1;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
#pragma unroll
        for (int j = 0; j < 5; ++j) v[rr][j] = rp[rr][min(x + j, xlast)];
    int nt[4], cls[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // corners 0..7 (marching_cubes.cu:118-127): rows (y,z), (y,z+1),
        // (y+1,z), (y+1,z+1) = rp[0..3], x at j, x+1 at j+1
        const float cv[8] = {v[0][j], v[0][j + 1], v[1][j + 1], v[1][j],
                             v[2][j], v[2][j + 1], v[3][j + 1], v[3][j]};
        cls[j] = cube_class(cv);
        nt[j] = (act && x + j < p.cx) ? (int)l.ntab[cls[j]] : 0;
    }
    const int c1 = nt[0], c2 = c1 + nt[1], c3 = c2 + nt[2];
    const int lt = c3 + nt[3];
    const int incl = wave_inclusive_scan(lt, lane);
    const int excl = incl - lt;
    const int gtotal = __builtin_amdgcn_readlane(incl, 63);
    const int cw = cls[0] | (cls[1] << 8) | (cls[2] << 16) | (cls[3] << 24);
    const int cp = c1 | (c2 << 8) | (c3 << 16);  // each <= 15
    __builtin_amdgcn_wave_barrier();
    // triangle k's owner: the last lane whose first triangle is <= k.  Each
    // round marks, at position (first - k0), the lanes whose first triangle
    // falls in [k0, k0 + 64) (lane + 1, 0 elsewhere); a max-scan of the marks
    // gives each position its owner, positions before the round's first mark
    // the previous round's last owner.
    int carry = 0;
    for (int k0 = 0; k0 < gtotal; k0 += 64) {
        const int k = k0 + lane;
        mark[lane] = 0;
        if (lt > 0 && excl >= k0 && excl < k0 + 64) mark[excl - k0] = lane + 1;
        __builtin_amdgcn_wave_barrier();
        const int mx = wave_inclusive_max(mark[lane]);
        const int owner = mx > 0 ? mx - 1 : carry;
        carry = __builtin_amdgcn_readlane(owner, 63);
        const int ocw = __shfl(cw, owner, 64);
        const int ocp = __shfl(cp, owner, 64);
        const int oex = __shfl(excl, owner, 64);
        const int ox = __shfl(x, owner, 64);
        const int orow = __shfl(r, owner, 64);
        const int local = k - oex;
        const int o1 = ocp & 0xFF, o2 = (ocp >> 8) & 0xFF, o3 = (ocp >> 16) & 0xFF;
        const int j = (local >= o1 ? 1 : 0) + (local >= o2 ? 1 : 0) + (local >= o3 ? 1 : 0);
        const int first = j == 0 ? 0 : (j == 1 ? o1 : (j == 2 ? o2 : o3));
        const int slot = local - first;
        const int oc = (ocw >> (8 * j)) & 0xFF;
        // the owner's voxel values: its cube j's corners are ov[..][j], ov[..][j+1]
        float ov[4][5];
#pragma unroll
        for (int rr = 0; rr < 4; ++rr)
#pragma unroll
            for (int q = 0; q < 5; ++q) ov[rr][q] = __shfl(v[rr][q], owner, 64);
        if (k < gtotal) {
            const int cx = ox + j, cy = orow % p.cyr, cz = orow / p.cyr;
            // load_corners' values (corners 0..7 as in the classification)
            float lo[4], hi[4];
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                lo[rr] = j == 0 ? ov[rr][0] : (j == 1 ? ov[rr][1] : (j == 2 ? ov[rr][2] : ov[rr][3]));
                hi[rr] = j == 0 ? ov[rr][1] : (j == 1 ? ov[rr][2] : (j == 2 ? ov[rr][3] : ov[rr][4]));
            }
            const float cv[8] = {lo[0], hi[0], hi[1], lo[1], lo[2], hi[2], hi[3], lo[3]};
            const uint64_t tw = l.tri[oc];
#pragma unroll
            for (int qv = 0; qv < 3; ++qv) {
                const int e = (int)((tw >> (4 + 4 * (3 * slot + qv))) & 0xFull);
                float pt[3];
                edge_vertex(p, cx, cy, cz, cv, e, pt);
                lstage[9 * lane + 3 * qv + 0] = pt[0];
                lstage[9 * lane + 3 * qv + 1] = pt[1];
                lstage[9 * lane + 3 * qv + 2] = pt[2];
            }
        }
        __builtin_amdgcn_wave_barrier();
        // triangles t0+k0 .. : 9 floats and 3 indices each, below capacity
        const long long tb = t0 + k0;
        const long long n = min((long long)min(64, gtotal - k0), capacity - tb);
        if (n > 0) {
            const int nw = 9 * (int)n;
            float* __restrict__ dst = verts + 9 * tb;
            for (int w = lane; w < nw; w += 64) dst[w] = lstage[w];
            if (tris) {
                int32_t* __restrict__ td = tris + 3 * tb;
                const int n3 = 3 * (int)n;
                for (int w = lane; w < n3; w += 64) td[w] = (int32_t)(3 * tb + w);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Tile b of a volume, whose triangles start at `base` (a 256-thread
// workgroup; every thread calls): its chunks with triangles, 256 at a time,
// are compacted in LDS (chunk index and first triangle in the tile, in
// order), then emitted four per wavefront.  l: the workgroup's LDS, its
// tables filled by mc_emit_lds_init (no barrier needed before the call: the
// first use follows one).
__device__ __forceinline__ void mc_emit_tile(const McParams& p, const McTiling& tl, const int b,
                                             const uint16_t* __restrict__ cc, const long long base,
                                             McEmitLds& l, float* __restrict__ verts,
                                             int32_t* __restrict__ tris, long long capacity) {
    const int gpl = tl.gpl, groups = tl.groups, nch = tl.nch;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int qa = b * kTileGroups, qb = min(qa + kTileGroups, groups) - 1;
    int ya, za, ka, yb, zb, kb;
    mc_group(p, gpl, qa, ya, za, ka);
    mc_group(p, gpl, qb, yb, zb, kb);
    const int c_begin = (ya + p.cyr * za) * nch;
    const int c_end = (yb + p.cyr * zb + kb) * nch;
    int woff = 0;  // triangles of the tile before the window
    for (int w0 = c_begin; w0 < c_end; w0 += 256) {
        const int ci = w0 + (int)threadIdx.x;
        const int cnt = ci < c_end ? (int)cc[ci] : 0;
        // packed: triangles (<= 256 * 320 < 2^20) and chunks with triangles
        const int pk = cnt | (cnt > 0 ? (1 << 20) : 0);
        const int incl = wave_inclusive_scan(pk, lane);
        if (lane == 63) l.wtot[wave] = incl;
        __syncthreads();
        int before = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) before += w < wave ? l.wtot[w] : 0;
        const int all = l.wtot[0] + l.wtot[1] + l.wtot[2] + l.wtot[3];
        const int ex = before + incl - pk;
        if (cnt > 0) {
            l.lchunk[ex >> 20] = ci;
            l.loff[ex >> 20] = woff + (ex & 0xFFFFF);
        }
        __syncthreads();
        const int nne = all >> 20;
        for (int g = 4 * wave; g < nne; g += 16)
            mc_emit_group(p, nch, l.lchunk + g, min(4, nne - g), base + l.loff[g], lane, l,
                          l.stage[wave], l.mark[wave], verts, tris, capacity);
        woff += all & 0xFFFFF;
        __syncthreads();
    }
}

// The last tile b in [lo, hi] with group_off[b] <= gi (group_off[lo] <= gi
// given), for one wavefront: each round probes 64 evenly spaced tiles at
// once, so 16384 tiles take three rounds of loads.
__device__ __forceinline__ int mc_find_tile(const int32_t* __restrict__ group_off, int lo, int hi,
                                            int gi, int lane) {
    while (hi > lo) {
        const int step = (hi - lo + 64) / 64;  // ceil((hi - lo + 1) / 64)
        const int idx = lo + lane * step;
        const bool le = idx <= hi && group_off[idx] <= gi;
        const unsigned long long m = __ballot(le);  // bit 0 is set
        const int last = 63 - __builtin_clzll(m);
        lo = lo + last * step;
        hi = min(hi, lo + step - 1);
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// Group gi of the volume for one wavefront (gi the same for every lane).
// Groups are runs of four consecutive chunks with triangles within one tile,
// tile by tile: group_off[b] is tile b's first group, group_off[ntiles]
// their total, tile_nne[b] its chunks with triangles.  The tile comes from
// mc_find_tile; the group's chunks and first triangles from a walk over the
// tile's chunk counts, 64 at a time (ballot of the chunks with triangles,
// prefix sums of their counts); then mc_emit_group.
__device__ __forceinline__ void mc_emit_group_at(const McParams& p, const McTiling& tl, int gi,
                                                 const uint16_t* __restrict__ cc,
                                                 const long long* __restrict__ tile_off,
                                                 const int32_t* __restrict__ tile_nne,
                                                 const int32_t* __restrict__ group_off,
                                                 McEmitLds& l, int wave, int lane,
                                                 float* __restrict__ verts,
                                                 int32_t* __restrict__ tris, long long capacity) {
    const int b = mc_find_tile(group_off, 0, tl.tiles - 1, gi, lane);
    const int qa = b * kTileGroups, qb = min(qa + kTileGroups, tl.groups) - 1;
    int ya, za, ka, yb, zb, kb;
    mc_group(p, tl.gpl, qa, ya, za, ka);
    mc_group(p, tl.gpl, qb, yb, zb, kb);
    const int c_begin = (ya + p.cyr * za) * tl.nch;
    const int c_end = (yb + p.cyr * zb + kb) * tl.nch;
    const int first = 4 * (gi - __builtin_amdgcn_readfirstlane(group_off[b]));
    const int want = min(4, __builtin_amdgcn_readfirstlane(tile_nne[b]) - first);
    int found = 0, ne_before = 0, tri_before = 0;
    for (int w0 = c_begin; w0 < c_end && found < want; w0 += 64) {
        const int ci = w0 + lane;
        const int cnt = ci < c_end ? (int)cc[ci] : 0;
        const unsigned long long m = __ballot(cnt > 0);
        const int rank = ne_before + (int)__popcll(m & ((1ull << lane) - 1ull));
        const int incl = wave_inclusive_scan(cnt, lane);
        if (cnt > 0 && rank >= first && rank < first + 4) {
            l.gchunk[wave][rank - first] = ci;
            l.goff[wave][rank - first] = tri_before + incl - cnt;
        }
        const int ne_here = (int)__popcll(m);
        found += max(0, min(first + 4, ne_before + ne_here) - max(first, ne_before));
        ne_before += ne_here;
        tri_before += __builtin_amdgcn_readlane(incl, 63);
    }
    __builtin_amdgcn_wave_barrier();
    mc_emit_group(p, tl.nch, l.gchunk[wave], want, tile_off[b] + l.goff[wave][0], lane, l,
                  l.stage[wave], l.mark[wave], verts, tris, capacity);
    __builtin_amdgcn_wave_barrier();
}

// Record of one group of four chunks with triangles (the emit's unit of
// work), written by mc_group_prep for the record-driven emit: bits 0-39 of
// tb its first triangle, bits 40-63 its tile; c[q] its q-th chunk counted
// from the tile's first chunk, kNoChunk past the last chunk of the tile's
// last group.  One 16-byte load gives a wavefront all of it.
struct alignas(16) McGroupRec {
    unsigned long long tb;
    uint16_t c[4];
};
constexpr int kRecTriBits = 40;
constexpr uint16_t kNoChunk = 0xFFFF;

// The first chunk of tile b and the end of its chunks (tiles are runs of
// kTileGroups consecutive groups, so their cube rows and chunks are
// consecutive).
__device__ __forceinline__ void mc_tile_chunks(const McParams& p, const McTiling& tl, int b,
                                               int& c_begin, int& c_end) {
    const int qa = b * kTileGroups, qb = min(qa + kTileGroups, tl.groups) - 1;
    int ya, za, ka, yb, zb, kb;
    mc_group(p, tl.gpl, qa, ya, za, ka);
    mc_group(p, tl.gpl, qb, yb, zb, kb);
    c_begin = (ya + p.cyr * za) * tl.nch;
    c_end = (yb + p.cyr * zb + kb) * tl.nch;
}

// The records of tile b's groups, for one wavefront: a walk over the tile's
// chunk counts, 256 at a time (four per lane; the lanes' counts of chunks
// with triangles and of triangles prefix-summed in one packed scan) gives
// each chunk with triangles its rank r and first triangle; it is slot r % 4
// of group group_off[b] + r / 4.  Every field of every record is written by
// exactly one lane.
__device__ __forceinline__ void mc_group_prep(const McParams& p, const McTiling& tl, int b,
                                              const uint16_t* __restrict__ cc,
                                              const long long* __restrict__ tile_off,
                                              const int32_t* __restrict__ tile_nne,
                                              const int32_t* __restrict__ group_off,
                                              McGroupRec* __restrict__ rec, int lane) {
    const int nne = tile_nne[b];
    if (nne == 0) return;
    const int g0 = group_off[b];
    const long long t0 = tile_off[b];
    int c_begin, c_end;
    mc_tile_chunks(p, tl, b, c_begin, c_end);
    const unsigned long long tile_bits = (unsigned long long)b << kRecTriBits;
    int ne_before = 0, tri_before = 0;
    for (int w0 = c_begin; w0 < c_end; w0 += 256) {
        const int ci0 = w0 + 4 * lane;
        int c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) c[q] = ci0 + q < c_end ? (int)cc[ci0 + q] : 0;
        int lane_ne = 0, lane_tri = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            lane_ne += c[q] > 0 ? 1 : 0;
            lane_tri += c[q];
        }
        // chunks with triangles (<= 256) in bits 0-8, triangles (<= 81920) above
        const int pk = lane_ne | (lane_tri << 9);
        const int incl = wave_inclusive_scan(pk, lane);
        int r = ne_before + ((incl - pk) & 0x1FF);
        int t = tri_before + ((incl - pk) >> 9);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (c[q] > 0) {
                McGroupRec& R = rec[g0 + (r >> 2)];
                R.c[r & 3] = (uint16_t)(ci0 + q - c_begin);
                if ((r & 3) == 0) R.tb = tile_bits | (unsigned long long)(t0 + t);
                if (r == nne - 1)
                    for (int e = (r & 3) + 1; e < 4; ++e) R.c[e] = kNoChunk;
                ++r;
            }
            t += c[q];
        }
        const int all = __builtin_amdgcn_readlane(incl, 63);
        ne_before += all & 0x1FF;
        tri_before += all >> 9;
    }
}

// Group gi from its record (mc_group_prep), for one wavefront; then
// mc_emit_group.
__device__ __forceinline__ void mc_emit_group_rec(const McParams& p, const McTiling& tl,
                                                  const McGroupRec& R, McEmitLds& l, int wave,
                                                  int lane, float* __restrict__ verts,
                                                  int32_t* __restrict__ tris, long long capacity) {
    const int b = (int)(R.tb >> kRecTriBits);
    const long long t0 = (long long)(R.tb & ((1ull << kRecTriBits) - 1ull));
    int c_begin, c_end;
    mc_tile_chunks(p, tl, b, c_begin, c_end);
    const int n = 1 + (R.c[1] != kNoChunk) + (R.c[2] != kNoChunk) + (R.c[3] != kNoChunk);
    if (lane < 4) l.gchunk[wave][lane] = c_begin + (lane == 0 ? R.c[0] : lane == 1 ? R.c[1]
                                                   : lane == 2 ? R.c[2] : R.c[3]);
    __builtin_amdgcn_wave_barrier();
    mc_emit_group(p, tl.nch, l.gchunk[wave], n, t0, lane, l, l.stage[wave], l.mark[wave], verts,
                  tris, capacity);
    __builtin_amdgcn_wave_barrier();
}

}  // namespace recon
