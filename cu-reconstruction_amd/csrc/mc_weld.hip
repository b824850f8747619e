// mc_weld.hip -- welded marching cubes for gfx950: shared vertices, per-vertex
// normals, and the soup's triangles as indices.
//
// No counterpart in the reference, whose generate_triangular_surface emits a
// soup (the per-triangle vert_remap of marching_cubes.cu:171-179 never shares
// a vertex).  Here one vertex per grid edge whose end values differ in sign
// (v < 0 against not), owned by its lower end voxel and numbered in the order
// (voxel x + gx*(y + gy*z), axis x, y, z); triangles are the soup's, in the
// soup's order (cube z, y, x, table slot).
//
// Numbering without a per-voxel table: the voxels of an x-row are cut into
// 64-voxel chunks, one wavefront each; three 64-bit ballots mark the chunk's
// sign-changing x, y and z edges and an exclusive scan over the chunk totals
// gives each chunk's first vertex.  The index of edge (voxel, axis) is then
//   base[chunk] + popc of the three masks below the voxel's lane
//               + the voxel's own set bits of the axes before `axis`,
// which the vertex writer and the triangle writer both evaluate: 48 B of
// masks, counts and bases per 64 voxels.  Two passes over the field, like
// the soup: weld_count_kernel (edge masks and the chunk's triangle count),
// two exclusive scans, weld_write_kernel (vertices, normals, triangles).
//
// Vertex: from the negative end N to the other end M, r = (-dN)/(dM - dN),
// p = sN + r*(sM - sN) per component.  This is the soup's own formula
// (marching_cubes.cu:23-51) whenever the soup orients the edge from N,
// i.e. except when dM is exactly +-0 and the cube lists M first (there the
// soup's bits already differ between the cubes sharing the edge).
// Normal: the TSDF gradient, central differences over (vs + vs) inside the
// grid and one-sided over vs on its border, lerped from N to M with r and
// normalised; (0,0,0) where the gradient vanishes.
#pragma clang fp contract(off)

#include "mc_device.hpp"

#include <cmath>

namespace recon {

constexpr int kWeldWaves = 4;  // one wavefront per voxel row, 256 threads

struct WeldParams {
    McParams mc;
    int32_t cpr;      // 64-voxel chunks per x-row
    int32_t nchunks;  // cpr * gy * gz
};

// Chunk c covers voxels x in [64k, 64k + 64) of voxel row (y, z), and the
// cubes with the same lower corners.  One pass marks the chunk's
// sign-changing edges and counts its cubes' triangles.
struct ChunkCorners {
    int x, y, z;
    float v[8];    // corner values of cube (x, y, z) (valid when `cube`)
    bool cube;     // (x, y, z) is the lower corner of a cube
    bool in;       // (x, y, z) is a voxel
};

__device__ __forceinline__ ChunkCorners chunk_corners(const WeldParams& w, int c, int lane) {
    const McParams& p = w.mc;
    ChunkCorners k;
    const int row = c / w.cpr;  // y + gy*z
    k.x = (c % w.cpr) * 64 + lane;
    k.y = row % p.gy;
    k.z = row / p.gy;
    k.in = k.x < p.gx;
    k.cube = k.x < p.gx - 1 && k.y < p.gy - 1 && k.z < p.gz - 1;
    if (k.cube) {
        load_corners(p, k.x, k.y, k.z, k.v);
    } else if (k.in) {
        // the voxel and whichever +x, +y, +z neighbours exist, in the
        // corner slots 0, 1, 4, 3 (all edge tests read only these)
        const int64_t i = (int64_t)k.x + (int64_t)p.gx * row;
        k.v[0] = p.f[i];
        k.v[1] = k.x + 1 < p.gx ? p.f[i + 1] : k.v[0];
        k.v[4] = k.y + 1 < p.gy ? p.f[i + p.gx] : k.v[0];
        k.v[3] = k.z + 1 < p.gz ? p.f[i + (int64_t)p.gx * p.gy] : k.v[0];
    }
    return k;
}

// Count pass: one wavefront per four voxel rows (y..y+3, z), streaming the
// ten voxel rows (y+i, z) and (y+i, z+1), i = 0..4, 64 voxels a step (the
// marching-cubes count's walk): each voxel's sign bit (v < 0) for the ten
// rows packs into one word, bit 2j the lane's voxel of row j and bit 2j+1
// the next lane's (a DPP wave shift; lane 63 takes the next step's lane 0).
// Per row i and chunk: the x, y and z edge masks are ballots of bit 4i
// against bits 4i+1, 4i+4 and 4i+2; the cube byte of cube row i is bits 4i
// .. 4i+7 (corner order 0, 1, 3, 2, 4, 5, 7, 6, the count table
// mc_ntab_qbyte); lane i of the wave stores row i's masks and counts.
// Loads past a row's end read its last voxel, and rows past the grid (y+i
// >= gy, z+1 >= gz) read the last row (in bounds); their edges and cubes
// do not count.
constexpr int kWeldRows = 4;
constexpr int kWeldLoadRows = 2 * (kWeldRows + 1);

__global__ __launch_bounds__(256) void weld_count_kernel(const WeldParams w,
                                                         unsigned long long* __restrict__ masks,
                                                         int32_t* __restrict__ vcount,
                                                         int32_t* __restrict__ tcount) {
    __shared__ uint8_t ntab[256];
    ntab[threadIdx.x] = mc_ntab_qbyte((int)threadIdx.x);
    __syncthreads();
    const McParams& p = w.mc;
    const int lane = threadIdx.x & 63;
    const int gpl = (p.gy + kWeldRows - 1) / kWeldRows;
    const int q = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kWeldWaves + (threadIdx.x >> 6)));
    const int z = q / gpl, y = kWeldRows * (q - z * gpl);
    if (z >= p.gz) return;
    const int k = min(kWeldRows, p.gy - y);         // voxel rows of the wave
    const int z1 = min(z + 1, p.gz - 1);
    const bool zin = z + 1 < p.gz;
    const char* rowp[kWeldLoadRows];
#pragma unroll
    for (int i = 0; i <= kWeldRows; ++i) {
        const int yy = min(y + i, p.gy - 1);
        rowp[2 * i] = reinterpret_cast<const char*>(p.f + (size_t)p.gx * ((size_t)yy + (size_t)p.gy * z));
        rowp[2 * i + 1] =
            reinterpret_cast<const char*>(p.f + (size_t)p.gx * ((size_t)yy + (size_t)p.gy * z1));
    }
    const uint32_t xlast = (uint32_t)(p.gx - 1);
    auto load = [&](int chunk, float v[kWeldLoadRows]) {
        const uint32_t off = min((uint32_t)(64 * chunk + lane), xlast) << 2;
#pragma unroll
        for (int j = 0; j < kWeldLoadRows; ++j) v[j] = *reinterpret_cast<const float*>(rowp[j] + off);
    };
    auto signs = [&](const float v[kWeldLoadRows]) {
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < kWeldLoadRows; ++j) m |= (v[j] < 0.0f ? 1u : 0u) << (2 * j);
        return m;
    };
    const int cpr = w.cpr;
    auto emit = [&](int c, uint32_t pc, uint32_t pn) {
        const uint32_t sh = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pc, 0x130, 0xF, 0xF, false);
        const uint32_t first = (uint32_t)__builtin_amdgcn_readlane((int)pn, 0);
        const uint32_t word = pc | ((lane == 63 ? first : sh) << 1);
        const int x = 64 * c + lane;
        const bool xin = x < p.gx, xin1 = x + 1 < p.gx;
        int n[kWeldRows];
        unsigned long long mx[kWeldRows], my[kWeldRows], mz[kWeldRows];
#pragma unroll
        for (int i = 0; i < kWeldRows; ++i) {
            const uint32_t s0 = (word >> (4 * i)) & 1u;
            const bool bx = xin1 && s0 != ((word >> (4 * i + 1)) & 1u);
            const bool by = xin && y + i + 1 < p.gy && s0 != ((word >> (4 * i + 4)) & 1u);
            const bool bz = xin && zin && s0 != ((word >> (4 * i + 2)) & 1u);
            mx[i] = __ballot(bx);
            my[i] = __ballot(by);
            mz[i] = __ballot(bz);
            const bool cube = xin1 && zin && y + i + 1 < p.gy;
            n[i] = cube ? (int)ntab[(word >> (4 * i)) & 0xFFu] : 0;
        }
        const int s01 = wave_inclusive_scan(n[0] | (n[1] << 16), lane);
        const int s23 = wave_inclusive_scan(n[2] | (n[3] << 16), lane);
        const int t01 = __builtin_amdgcn_readlane(s01, 63), t23 = __builtin_amdgcn_readlane(s23, 63);
        if (lane < k) {
            const int i = lane;
            const unsigned long long ax = i == 0 ? mx[0] : (i == 1 ? mx[1] : (i == 2 ? mx[2] : mx[3]));
            const unsigned long long ay = i == 0 ? my[0] : (i == 1 ? my[1] : (i == 2 ? my[2] : my[3]));
            const unsigned long long az = i == 0 ? mz[0] : (i == 1 ? mz[1] : (i == 2 ? mz[2] : mz[3]));
            const int tw = i < 2 ? t01 : t23;
            const int64_t cc = (int64_t)(y + i + p.gy * z) * cpr + c;
            masks[3 * cc + 0] = ax;
            masks[3 * cc + 1] = ay;
            masks[3 * cc + 2] = az;
            vcount[cc] = __popcll(ax) + __popcll(ay) + __popcll(az);
            tcount[cc] = (i & 1) ? (tw >> 16) : (tw & 0xFFFF);
        }
    };
    float va[kWeldLoadRows], vb[kWeldLoadRows];
    load(0, va);
    load(1, vb);
    uint32_t pc = signs(va);
    load(2, va);
    for (int c = 0; c < cpr; c += 2) {
        uint32_t pn = signs(vb);  // chunk c+1
        load(c + 3, vb);
        emit(c, pc, pn);
        if (c + 1 >= cpr) break;
        pc = pn;
        pn = signs(va);  // chunk c+2
        load(c + 4, va);
        emit(c + 1, pc, pn);
        pc = pn;
    }
}

// Gradient component `a` of the field at voxel (x, y, z).
__device__ __forceinline__ float weld_grad(const McParams& p, int x, int y, int z, int a) {
    const int64_t st = a == 0 ? 1 : (a == 1 ? (int64_t)p.gx : (int64_t)p.gx * p.gy);
    const int c = a == 0 ? x : (a == 1 ? y : z);
    const int g = a == 0 ? p.gx : (a == 1 ? p.gy : p.gz);
    const int64_t i = (int64_t)x + (int64_t)p.gx * (y + (int64_t)p.gy * z);
    if (c > 0 && c < g - 1) return (p.f[i + st] - p.f[i - st]) / (p.vs[a] + p.vs[a]);
    if (c == 0) return (p.f[i + st] - p.f[i]) / p.vs[a];
    return (p.f[i] - p.f[i - st]) / p.vs[a];
}

// Vertex (and normal) `idx` on the edge from voxel (x, y, z) along axis a.
__device__ __forceinline__ void weld_vertex(const WeldParams& w, int x, int y, int z, int a,
                                            long long idx, float* __restrict__ verts,
                                            float* __restrict__ normals) {
    const McParams& p = w.mc;
    const int row = y + p.gy * z;
    const int q[3] = {x + (a == 0), y + (a == 1), z + (a == 2)};
    const int64_t i = (int64_t)x + (int64_t)p.gx * row;
    const int64_t j = (int64_t)q[0] + (int64_t)p.gx * (q[1] + (int64_t)p.gy * q[2]);
    const bool neg_first = p.f[i] < 0.0f;
    const int pn[3] = {neg_first ? x : q[0], neg_first ? y : q[1], neg_first ? z : q[2]};
    const int pm[3] = {neg_first ? q[0] : x, neg_first ? q[1] : y, neg_first ? q[2] : z};
    const float dn = neg_first ? p.f[i] : p.f[j];
    const float dm = neg_first ? p.f[j] : p.f[i];
    const float r = (-dn) / (dm - dn);
    if (verts) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float sn = voxel_centre(pn[k], p.vs[k], p.off[k]);
            const float sm = voxel_centre(pm[k], p.vs[k], p.off[k]);
            const float e = sm - sn;
            const float o = r * e;
            verts[3 * idx + k] = sn + o;
        }
    }
    if (normals) {
        float g[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float gn = weld_grad(p, pn[0], pn[1], pn[2], k);
            const float gm = weld_grad(p, pm[0], pm[1], pm[2], k);
            const float e = gm - gn;
            const float o = r * e;
            g[k] = gn + o;
        }
        const float len2 = (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2];
        if (len2 > 0.0f) {
            const float len = sqrtf(len2);
#pragma unroll
            for (int k = 0; k < 3; ++k) normals[3 * idx + k] = g[k] / len;
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) normals[3 * idx + k] = 0.0f;
        }
    }
}

// The vertices (and normals) of chunk c, numbered base + k in the order lane,
// then axis: vertex k is computed by lane k % 64 (a lane owning three edges
// does not hold up the wave).  Every lane must call.
__device__ __forceinline__ void weld_chunk_vertices(const WeldParams& w, int c, int lane,
                                                    const unsigned long long m[3], long long base,
                                                    float* __restrict__ verts,
                                                    float* __restrict__ normals) {
    const McParams& p = w.mc;
    const int row = c / w.cpr;
    const int x0 = (c % w.cpr) * 64;
    const int y = row % p.gy, z = row / p.gy;
    const int b0 = (int)((m[0] >> lane) & 1ull), b1 = (int)((m[1] >> lane) & 1ull),
              b2 = (int)((m[2] >> lane) & 1ull);
    const int cnt = b0 + b1 + b2;
    const int incl = wave_inclusive_scan(cnt, lane);
    const int excl = incl - cnt;
    const int total = __popcll(m[0]) + __popcll(m[1]) + __popcll(m[2]);
    for (int k0 = 0; k0 < total; k0 += 64) {
        const int k = k0 + lane;
        const int owner = wave_owner(incl, k);
        const int r = k - __shfl(excl, owner, 64);  // rank among the owner's set axes
        const int o0 = (int)((m[0] >> owner) & 1ull), o1 = (int)((m[1] >> owner) & 1ull);
        const int a = r == 0 ? (o0 ? 0 : (o1 ? 1 : 2)) : (r == 1 ? (o0 && o1 ? 1 : 2) : 2);
        if (k < total) weld_vertex(w, x0 + owner, y, z, a, base + k, verts, normals);
    }
}


// The chunk's vertices, then its cubes' triangles in the soup's order (one
// wavefront; every lane calls).
// Masks and first vertices of the 8 chunks a chunk's cubes reach (rows y,
// y+1 by layers z, z+1, by x-chunks j, j+1), per wavefront in LDS, so that a
// triangle's vertex indices need no global loads: neighbour n = 4 dy + 2 dz
// + dj; chunks past the grid hold zeros (no cube reaches them).
struct WeldNbr {
    unsigned long long m[8][3];
    long long base[8];
};

__device__ __forceinline__ void weld_load_nbr(const WeldParams& w, int c, int lane,
                                              const unsigned long long* __restrict__ masks,
                                              const long long* __restrict__ vbase, WeldNbr& nb) {
    if (lane < 8) {
        const int row = c / w.cpr, j = c - row * w.cpr;
        const int y = row % w.mc.gy, z = row / w.mc.gy;
        const int Y = y + ((lane >> 2) & 1), Z = z + ((lane >> 1) & 1), J = j + (lane & 1);
        unsigned long long m0 = 0, m1 = 0, m2 = 0;
        long long b = 0;
        if (Y < w.mc.gy && Z < w.mc.gz && J < w.cpr) {
            const int64_t cn = (int64_t)(Y + w.mc.gy * Z) * w.cpr + J;
            m0 = masks[3 * cn];
            m1 = masks[3 * cn + 1];
            m2 = masks[3 * cn + 2];
            b = vbase[cn];
        }
        nb.m[lane][0] = m0;
        nb.m[lane][1] = m1;
        nb.m[lane][2] = m2;
        nb.base[lane] = b;
    }
    __builtin_amdgcn_wave_barrier();
}

// Vertex index of local edge e of cube (x, y, z) of the chunk whose
// neighbours nb holds (x0 the chunk's first voxel): the base of the edge's
// chunk, plus the set bits of the three masks below the edge's lower end,
// plus that voxel's own bits of the axes before the edge's.
__device__ __forceinline__ int32_t weld_index_nbr(const WeldNbr& nb, int x, int x0, int e) {
    // corner offsets packed as bits: dx = bit0, dy = bit1, dz = bit2
    // (0:000 1:001 2:101 3:100 4:010 5:011 6:111 7:110), as in edge_vertex
    constexpr unsigned k_off = (0u << 0) | (1u << 3) | (5u << 6) | (4u << 9) | (2u << 12) |
                               (3u << 15) | (7u << 18) | (6u << 21);
    const int ca = (int)((kEdgeA >> (4 * e)) & 0xFull), cb = (int)((kEdgeB >> (4 * e)) & 0xFull);
    const unsigned oa = (k_off >> (3 * ca)) & 7u, ob = (k_off >> (3 * cb)) & 7u;
    const int a = __builtin_ctz(oa ^ ob);  // the axis the edge runs along
    const unsigned lo = oa & ob;           // its lower end
    const int X = x + (int)(lo & 1u);
    const int n = (int)(((lo >> 1) & 1u) << 2) | (int)((lo >> 2) << 1) | ((X >> 6) - (x0 >> 6));
    const int l = X & 63;
    const unsigned long long m0 = nb.m[n][0], m1 = nb.m[n][1], m2 = nb.m[n][2];
    const unsigned long long below = (1ull << l) - 1ull;
    long long idx = nb.base[n] + __popcll(m0 & below) + __popcll(m1 & below) + __popcll(m2 & below);
    if (a > 0) idx += (long long)((m0 >> l) & 1ull);
    if (a > 1) idx += (long long)((m1 >> l) & 1ull);
    return (int32_t)idx;
}

__device__ __forceinline__ void weld_write_chunk(const WeldParams& w, int c, int lane,
                                                 const unsigned long long* __restrict__ masks,
                                                 const long long* __restrict__ vbase,
                                                 const long long* __restrict__ tbase, int nt_chunk,
                                                 float* __restrict__ verts,
                                                 float* __restrict__ normals,
                                                 int32_t* __restrict__ tris, WeldNbr& nb) {
    if (verts || normals) {
        const unsigned long long m[3] = {masks[3 * (int64_t)c], masks[3 * (int64_t)c + 1],
                                         masks[3 * (int64_t)c + 2]};
        if (m[0] | m[1] | m[2]) weld_chunk_vertices(w, c, lane, m, vbase[c], verts, normals);
    }
    if (!tris || nt_chunk == 0) return;
    weld_load_nbr(w, c, lane, masks, vbase, nb);
    const ChunkCorners k = chunk_corners(w, c, lane);
    int cls = 0, nt = 0;
    if (k.cube) {
        cls = cube_class(k.v);
        nt = class_ntri(cls);
    }
    const int incl = wave_inclusive_scan(nt, lane);
    const int excl = incl - nt;
    const int total = __shfl(incl, 63, 64);
    const long long t0 = tbase[c];
    const int x0 = k.x - lane;
    // triangle kk of the chunk is written by lane kk % 64 (its cube found by
    // wave_owner over the lanes' prefix counts)
    for (int k0 = 0; k0 < total; k0 += 64) {
        const int kk = k0 + lane;
        const int owner = wave_owner(incl, kk);
        const int ocls = __shfl(cls, owner, 64);
        const int slot = kk - __shfl(excl, owner, 64);
        if (kk < total) {
            const long long t = t0 + kk;
            const uint64_t word = c_tri[ocls];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int e = (int)((word >> (4 + 4 * (3 * slot + q))) & 0xFull);
                tris[3 * t + q] = weld_index_nbr(nb, x0 + owner, x0, e);
            }
        }
    }
}

// One wavefront per voxel row: one load of the vertex and triangle counts
// of 64 of the row's chunks at a time, a ballot of the chunks with either,
// and weld_write_chunk for each of those (most of a volume's chunks hold
// neither, and an empty row costs one load).  Nothing is written when the
// totals (*vtotal, *ttotal, from the scans) exceed int32 indices or a
// requested output's capacity: the host reports that after the call.
__global__ __launch_bounds__(256) void weld_write_kernel(const WeldParams w,
                                                         const unsigned long long* __restrict__ masks,
                                                         const int32_t* __restrict__ vcount,
                                                         const long long* __restrict__ vbase,
                                                         const int32_t* __restrict__ tcount,
                                                         const long long* __restrict__ tbase,
                                                         const long long* __restrict__ vtotal,
                                                         const long long* __restrict__ ttotal,
                                                         long long vcap, long long tcap,
                                                         float* __restrict__ verts,
                                                         float* __restrict__ normals,
                                                         int32_t* __restrict__ tris) {
    __shared__ WeldNbr nbr[kWeldWaves];
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * kWeldWaves + (threadIdx.x >> 6);
    if (row >= w.mc.gy * w.mc.gz) return;
    WeldNbr& nb = nbr[threadIdx.x >> 6];
    const bool want_v = verts || normals;
    const long long nvt = *vtotal, ntt = *ttotal;
    if (nvt > INT32_MAX || (want_v && nvt > vcap) || (tris && ntt > tcap)) return;
    const int c0 = row * w.cpr, c1 = c0 + w.cpr;
    for (int cw = c0; cw < c1; cw += 64) {
        const int c = cw + lane;
        const int nv = c < c1 && want_v ? vcount[c] : 0;
        const int nt = c < c1 && tris ? tcount[c] : 0;
        unsigned long long busy = __ballot(nv > 0 || nt > 0);
        while (busy) {
            const int l = __builtin_ctzll(busy);
            busy &= busy - 1;
            weld_write_chunk(w, cw + l, lane, masks, vbase, tbase, __shfl(nt, l, 64), verts,
                             normals, tris, nb);
        }
    }
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }


// Counts (always), then writes what is asked for when it fits.
static int weld_run(const float* field, const int32_t grid[3], const float vs[3], const float off[3],
                    float* verts, float* normals, int64_t vcap, int32_t* tris, int64_t tcap,
                    int64_t* nv_out, int64_t* nt_out, void** scratch, size_t* scratch_bytes,
                    hipStream_t s, const char* fn) {
    *nv_out = 0;
    *nt_out = 0;
    if (grid[0] < 2 || grid[1] < 2 || grid[2] < 2) return RECON_OK;
    WeldParams w;
    McParams& p = w.mc;
    p.f = field;
    p.gx = grid[0];
    p.gy = grid[1];
    p.gz = grid[2];
    p.cx = grid[0] - 1;
    p.cyr = grid[1] - 1;
    p.rows = (grid[1] - 1) * (grid[2] - 1);
    for (int i = 0; i < 3; ++i) {
        p.vs[i] = vs[i];
        p.off[i] = off[i];
    }
    w.cpr = (grid[0] + 63) / 64;
    w.nchunks = w.cpr * grid[1] * grid[2];
    const int nb = (w.nchunks + kScanBlock - 1) / kScanBlock;
    const size_t o_masks = 0;
    const size_t o_vcount = o_masks + align256(24 * (size_t)w.nchunks);
    const size_t o_tcount = o_vcount + align256(4 * (size_t)w.nchunks);
    const size_t o_vbase = o_tcount + align256(4 * (size_t)w.nchunks);
    const size_t o_tbase = o_vbase + align256(8 * (size_t)w.nchunks);
    const size_t o_vsums = o_tbase + align256(8 * (size_t)w.nchunks);
    const size_t o_tsums = o_vsums + align256(8 * (size_t)(nb + 1));
    const size_t need = o_tsums + align256(8 * (size_t)(nb + 1));
    if (*scratch_bytes < need) {
        if (*scratch) (void)hipFree(*scratch);
        *scratch = nullptr;
        *scratch_bytes = 0;
        RECON_HIP_TRY(hipMalloc(scratch, need), "welded marching cubes scratch");
        *scratch_bytes = need;
    }
    char* sb = static_cast<char*>(*scratch);
    auto* masks = reinterpret_cast<unsigned long long*>(sb + o_masks);
    auto* vcount = reinterpret_cast<int32_t*>(sb + o_vcount);
    auto* tcount = reinterpret_cast<int32_t*>(sb + o_tcount);
    auto* vbase = reinterpret_cast<long long*>(sb + o_vbase);
    auto* tbase = reinterpret_cast<long long*>(sb + o_tbase);
    auto* vsums = reinterpret_cast<long long*>(sb + o_vsums);
    auto* tsums = reinterpret_cast<long long*>(sb + o_tsums);
    const int count_blocks =
        (((grid[1] + kWeldRows - 1) / kWeldRows) * grid[2] + kWeldWaves - 1) / kWeldWaves;
    hipLaunchKernelGGL(weld_count_kernel, dim3(count_blocks), dim3(256), 0, s, w, masks, vcount,
                       tcount);
    RECON_HIP_TRY(hipGetLastError(), "weld_count_kernel launch");
    ScanJobs jobs;
    jobs.in[0] = vcount;
    jobs.in[1] = tcount;
    jobs.sums[0] = vsums;
    jobs.sums[1] = tsums;
    jobs.out[0] = vbase;
    jobs.out[1] = tbase;
    const int rc = launch_exclusive_scans(jobs, 2, w.nchunks, s);
    if (rc) return rc;
    // The write pass checks the totals on the device, so it is queued before
    // the host waits; the copies of the totals to pageable memory come last,
    // since each returns only when the stream has reached it.
    const bool write = verts || normals || tris;
    if (write) {
        const int blocks = (grid[1] * grid[2] + kWeldWaves - 1) / kWeldWaves;
        hipLaunchKernelGGL(weld_write_kernel, dim3(blocks), dim3(256), 0, s, w, masks, vcount,
                           vbase, tcount, tbase, vsums + nb, tsums + nb, (long long)vcap,
                           (long long)tcap, verts, normals, tris);
        RECON_HIP_TRY(hipGetLastError(), "weld_write_kernel launch");
    }
    long long totals[2] = {0, 0};
    RECON_HIP_TRY(hipMemcpyAsync(&totals[0], vsums + nb, 8, hipMemcpyDeviceToHost, s), "vertex total");
    RECON_HIP_TRY(hipMemcpyAsync(&totals[1], tsums + nb, 8, hipMemcpyDeviceToHost, s), "triangle total");
    RECON_HIP_TRY(hipStreamSynchronize(s), "welded marching cubes");
    *nv_out = totals[0];
    *nt_out = totals[1];
    if (!write) return RECON_OK;
    if (totals[0] > INT32_MAX)
        return fail(RECON_ERR_RANGE, "%s: %lld vertices exceed int32 indices", fn, totals[0]);
    if ((verts || normals) && totals[0] > vcap)
        return fail(RECON_ERR_CAPACITY, "%s: %lld vertices do not fit capacity %lld", fn, totals[0],
                    (long long)vcap);
    if (tris && totals[1] > tcap)
        return fail(RECON_ERR_CAPACITY, "%s: %lld triangles do not fit capacity %lld", fn, totals[1],
                    (long long)tcap);
    return RECON_OK;
}

static int weld_entry(const float* field, const int32_t grid[3], const float vs[3],
                      const float off[3], float* verts, float* normals, int64_t vcap,
                      int32_t* tris, int64_t tcap, int64_t* nv, int64_t* nt, void** scratch,
                      size_t* scratch_bytes, hipStream_t s, const char* fn) {
    if (!nv || !nt) return fail(RECON_ERR_NULL, "%s: count output is NULL", fn);
    *nv = 0;
    *nt = 0;
    if (!field || !grid || !vs || !off) return fail(RECON_ERR_NULL, "%s: null argument", fn);
    int64_t n = 1;
    for (int i = 0; i < 3; ++i) {
        if (grid[i] < 1) return fail(RECON_ERR_SIZE, "%s: grid[%d] = %d must be >= 1", fn, i, grid[i]);
        n *= grid[i];
    }
    if (n >= (int64_t(1) << 31)) return fail(RECON_ERR_RANGE, "%s: grid exceeds 2^31-1 voxels", fn);
    if (vcap < 0 || tcap < 0) return fail(RECON_ERR_SIZE, "%s: negative capacity", fn);
    const int rc = require_gfx950();
    if (rc) return rc;
    return weld_run(field, grid, vs, off, verts, normals, vcap, tris, tcap, nv, nt, scratch,
                    scratch_bytes, s, fn);
}

}  // namespace recon

using namespace recon;

extern "C" int recon_mc_extract_welded(const float* field, const int32_t grid[3],
                                       const float voxel_size[3], const float offset[3],
                                       float* vertices, float* normals, int64_t vertex_capacity,
                                       int32_t* triangles, int64_t triangle_capacity,
                                       int64_t* num_vertices, int64_t* num_triangles,
                                       recon_stream_t stream) {
    clear_error();
    void* scratch = nullptr;
    size_t bytes = 0;
    const int rc = weld_entry(field, grid, voxel_size, offset, vertices, normals, vertex_capacity,
                              triangles, triangle_capacity, num_vertices, num_triangles, &scratch,
                              &bytes, (hipStream_t)stream, "recon_mc_extract_welded");
    if (scratch) (void)hipFree(scratch);
    return rc;
}

extern "C" int recon_tsdf_extract_welded_mesh(recon_tsdf* h, float* vertices, float* normals,
                                              int64_t vertex_capacity, int32_t* triangles,
                                              int64_t triangle_capacity, int64_t* num_vertices,
                                              int64_t* num_triangles, recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_tsdf_extract_welded_mesh";
    if (!h) {
        if (num_vertices) *num_vertices = 0;
        if (num_triangles) *num_triangles = 0;
        return fail(RECON_ERR_NULL, "%s: handle is NULL", fn);
    }
    return weld_entry(h->D, h->info.grid, h->info.voxel_size, h->info.global_offset, vertices,
                      normals, vertex_capacity, triangles, triangle_capacity, num_vertices,
                      num_triangles, &h->mc_scratch, &h->mc_scratch_bytes, (hipStream_t)stream, fn);
}
