// marching_cubes.hip -- marching cubes over a dense TSDF for gfx950.
//
// Replaces generate_triangular_surface (cu-Reconstruction
// modules/src/cuda/marching_cubes.cu:102-236).  The reference launches one
// kernel per z layer, synchronises, and compacts each layer on the host from
// managed memory.  Here the whole volume is done on the device in three
// launches with a fixed, deterministic output order:
//   1. chunk count - one wavefront per 4 cube rows (y..y+3, z), streaming
//                    their 10 voxel rows once, 64 cubes a step: every
//                    cube's class (marching_cubes.cu:76-84,118-130) from
//                    the corners' sign bits, its triangle count from a
//                    table, and per 64-cube chunk of each row the chunk's
//                    triangle count; per workgroup ("tile", 16 cube rows)
//                    the tile's total;
//   2. tile scan   - one workgroup: exclusive scan of the tile totals, each
//                    tile's first triangle and the volume's total;
//   3. emit        - one workgroup per tile: the tile's chunks with
//                    triangles are compacted in LDS and taken four at a
//                    time by a wavefront (16 lanes per chunk, 4 cubes per
//                    lane), whose triangles are spread over its 64 lanes
//                    to compute and write the vertices
//                    (marching_cubes.cu:23-74).
// Output order is cube z, y, x, table slot -- the order the reference's host
// loop produces (marching_cubes.cu:163-188 inside :213-225) -- as a triangle
// soup: triangle t owns vertices 3t, 3t+1, 3t+2 (marching_cubes.cu:171).
#pragma clang fp contract(off)

#include "mc_device.hpp"

#include <algorithm>
#include <mutex>

namespace recon {

constexpr int kRowsPerBlock = 4;  // one wavefront per row, 256 threads

// ---- chunk count --------------------------------------------------------------
// A wavefront takes a group of kGroupRows consecutive cube rows of one layer
// (the layer's last group may be shorter); a workgroup (tile) takes
// kTileGroups consecutive groups, so a tile covers consecutive cube rows.
constexpr int kGroupRows = 4;
constexpr int kTileGroups = 4;
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

// Sign bits (v < 0, so -0 and NaN give 0: define_voxel_type,
// marching_cubes.cu:76-84) of the kVoxelRows values, row j at bit 2j.
__device__ __forceinline__ uint32_t mc_sign_bits(const float v[kVoxelRows]) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < kVoxelRows; ++j) m |= (v[j] < 0.0f ? 1u : 0u) << (2 * j);
    return m;
}

// Triangles of each 64-cube chunk of every cube row into cc (chunk c of row
// r at r*nch + c), and each tile's total into tile_sum.  The chunk's voxel
// rows are loaded two chunks ahead; a cube's x+1 corners are the next lane's
// sign bits (a DPP wave shift), lane 63's from the following chunk.  Loads
// past the row end read its last voxel, and rows past the layer's last
// voxel row read that row (both in bounds); their cubes are not counted.
__global__ __launch_bounds__(256) void mc_chunk_count_kernel(const McParams p, const int gpl,
                                                             const int groups, const int nch,
                                                             uint16_t* __restrict__ cc,
                                                             int32_t* __restrict__ tile_sum) {
    __shared__ uint8_t ntab[256];
    __shared__ int wtot[4];
    ntab[threadIdx.x] = (uint8_t)class_ntri(qbyte_class((int)threadIdx.x));
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int q = blockIdx.x * kTileGroups + wave;
    int total = 0;
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
                ccr[(size_t)lane * nch + c] = (uint16_t)((lane & 1) ? (w >> 16) : (w & 0xFFFF));
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
    if (lane == 0) wtot[threadIdx.x >> 6] = total;
    __syncthreads();
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = wtot[0] + wtot[1] + wtot[2] + wtot[3];
}

// ---- tile scan ------------------------------------------------------------------
constexpr int kTileScanItems = 16;

// Inclusive prefix sum of a 64-bit value over the lanes of a wavefront.
__device__ __forceinline__ long long wave_inclusive_scan64(long long v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const long long o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// One workgroup of 1024 threads: tile_off[i] = sum of tile_sum[0..i), and
// tile_off[n] = the total.
__global__ __launch_bounds__(1024) void mc_tile_scan_kernel(const int32_t* __restrict__ tile_sum,
                                                            int n, long long* __restrict__ tile_off) {
    __shared__ long long wtot[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    long long carry = 0;
    for (int base = 0; base < n; base += 1024 * kTileScanItems) {
        const int i0 = base + (int)threadIdx.x * kTileScanItems;
        int v[kTileScanItems];
        long long sum = 0;
#pragma unroll
        for (int j = 0; j < kTileScanItems; ++j) {
            v[j] = i0 + j < n ? tile_sum[i0 + j] : 0;
            sum += v[j];
        }
        const long long incl = wave_inclusive_scan64(sum, lane);
        if (lane == 63) wtot[wave] = incl;
        __syncthreads();
        long long before = carry, all = 0;
#pragma unroll
        for (int w = 0; w < 16; ++w) {
            before += w < wave ? wtot[w] : 0;
            all += wtot[w];
        }
        long long ex = before + incl - sum;
#pragma unroll
        for (int j = 0; j < kTileScanItems; ++j) {
            if (i0 + j < n) tile_off[i0 + j] = ex;
            ex += v[j];
        }
        carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_off[n] = carry;
}

// ---- emit -----------------------------------------------------------------------
// Writes the triangles of up to four chunks (list entries e0 .. e0+ne-1 of
// the tile's compacted list, consecutive in the output from triangle t0) by
// one wavefront: lane l takes cubes 4(l%16) .. +3 of chunk l/16, so the
// lanes' cubes are in output order; a wave-level prefix sum places each
// lane's triangles, and triangle k of the group is written by lane k % 64,
// which finds its cube (wave_owner, then among the owner's four), reloads
// the cube's corners and computes the vertices (edge_vertex).
__device__ __forceinline__ void mc_emit_group(const McParams& p, int nch, const int* lchunk,
                                              int ne, long long t0, int lane,
                                              const uint8_t* ntab, float* __restrict__ verts,
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
        nt[j] = (act && x + j < p.cx) ? (int)ntab[cls[j]] : 0;
    }
    const int c1 = nt[0], c2 = c1 + nt[1], c3 = c2 + nt[2];
    const int lt = c3 + nt[3];
    const int incl = wave_inclusive_scan(lt, lane);
    const int excl = incl - lt;
    const int gtotal = __builtin_amdgcn_readlane(incl, 63);
    const int cw = cls[0] | (cls[1] << 8) | (cls[2] << 16) | (cls[3] << 24);
    const int cp = c1 | (c2 << 8) | (c3 << 16);  // each <= 15
    for (int k0 = 0; k0 < gtotal; k0 += 64) {
        const int k = k0 + lane;
        const int owner = wave_owner(incl, k);
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
        const long long t = t0 + k;
        if (k < gtotal && t < capacity) {
            const int cx = ox + j, cy = orow % p.cyr, cz = orow / p.cyr;
            float cv[8];
            load_corners(p, cx, cy, cz, cv);
            const uint64_t tw = c_tri[oc];
#pragma unroll
            for (int qv = 0; qv < 3; ++qv) {
                const int e = (int)((tw >> (4 + 4 * (3 * slot + qv))) & 0xFull);
                float pt[3];
                edge_vertex(p, cx, cy, cz, cv, e, pt);
                float* dst = verts + 9 * t + 3 * qv;
                dst[0] = pt[0];
                dst[1] = pt[1];
                dst[2] = pt[2];
                if (tris) tris[3 * t + qv] = (int32_t)(3 * t + qv);
            }
        }
    }
}

// One workgroup per tile: its chunks with triangles, 256 at a time, are
// compacted in LDS (chunk index and first triangle in the tile, in order),
// then emitted four per wavefront.
__global__ __launch_bounds__(256) void mc_chunk_emit_kernel(const McParams p, const int gpl,
                                                            const int groups, const int nch,
                                                            const uint16_t* __restrict__ cc,
                                                            const long long* __restrict__ tile_off,
                                                            float* __restrict__ verts,
                                                            int32_t* __restrict__ tris,
                                                            long long capacity) {
    __shared__ uint8_t ntab[256];
    __shared__ int lchunk[256];
    __shared__ int loff[256];
    __shared__ int wtot[4];
    const int b = blockIdx.x;
    const long long base = tile_off[b];
    if (tile_off[b + 1] == base) return;  // the same for every thread
    ntab[threadIdx.x] = (uint8_t)class_ntri((int)threadIdx.x);
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
        if (lane == 63) wtot[wave] = incl;
        __syncthreads();
        int before = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) before += w < wave ? wtot[w] : 0;
        const int all = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        const int ex = before + incl - pk;
        if (cnt > 0) {
            lchunk[ex >> 20] = ci;
            loff[ex >> 20] = woff + (ex & 0xFFFFF);
        }
        __syncthreads();
        const int nne = all >> 20;
        for (int g = 4 * wave; g < nne; g += 16)
            mc_emit_group(p, nch, lchunk + g, min(4, nne - g), base + loff[g], lane, ntab, verts,
                          tris, capacity);
        woff += all & 0xFFFFF;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void mc_class_hist_kernel(const McParams p,
                                                            unsigned long long* __restrict__ hist) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    if (r < p.rows) {
        const int y = r % p.cyr, z = r / p.cyr;
        for (int x = lane; x < p.cx; x += 64) {
            float v[8];
            load_corners(p, x, y, z, v);
            atomicAdd(&h[cube_class(v)], 1u);
        }
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void scan_block_sums_kernel(const int32_t* __restrict__ in, int n,
                                                              long long* __restrict__ sums) {
    __shared__ long long sh[256];
    const int b0 = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (b0 + k < n) s += in[b0 + k];
    long long total;
    block_exclusive_scan(s, sh, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the block sums in place, total at [nb].
__global__ __launch_bounds__(256) void scan_top_kernel(long long* __restrict__ sums, int nb) {
    __shared__ long long sh[256];
    long long carry = 0;
    for (int base = 0; base < nb; base += 256) {
        const int i = base + threadIdx.x;
        const long long v = i < nb ? sums[i] : 0;
        long long total;
        const long long ex = block_exclusive_scan(v, sh, &total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ __launch_bounds__(256) void scan_apply_kernel(const int32_t* __restrict__ in, int n,
                                                         const long long* __restrict__ sums,
                                                         long long* __restrict__ out) {
    __shared__ long long sh[256];
    const int b0 = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    long long vals[kScanItems];
    long long s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        vals[k] = (b0 + k < n) ? in[b0 + k] : 0;
        s += vals[k];
    }
    long long total;
    long long ex = block_exclusive_scan(s, sh, &total) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (b0 + k < n) out[b0 + k] = ex;
        ex += vals[k];
    }
}

int launch_exclusive_scan(const int32_t* counts, int n, long long* offsets, long long* sums,
                          hipStream_t s) {
    const int nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(scan_block_sums_kernel, dim3(nb), dim3(256), 0, s, counts, n, sums);
    RECON_HIP_TRY(hipGetLastError(), "scan_block_sums_kernel launch");
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(256), 0, s, sums, nb);
    RECON_HIP_TRY(hipGetLastError(), "scan_top_kernel launch");
    if (offsets) {
        hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(256), 0, s, counts, n, sums, offsets);
        RECON_HIP_TRY(hipGetLastError(), "scan_apply_kernel launch");
    }
    return RECON_OK;
}

static McParams make_params(const float* field, const int32_t grid[3], const float vs[3],
                            const float off[3]) {
    McParams p;
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
    return p;
}

static int check_grid(const float* field, const int32_t grid[3], const float vs[3],
                      const float off[3], const char* fn) {
    if (!field || !grid || !vs || !off) return fail(RECON_ERR_NULL, "%s: null argument", fn);
    int64_t n = 1;
    for (int i = 0; i < 3; ++i) {
        if (grid[i] < 1) return fail(RECON_ERR_SIZE, "%s: grid[%d] = %d must be >= 1", fn, i, grid[i]);
        n *= grid[i];
    }
    if (n >= (int64_t(1) << 31)) return fail(RECON_ERR_RANGE, "%s: grid exceeds 2^31-1 voxels", fn);
    return RECON_OK;
}

// Tiling of a grid (host side) and the scratch layout: chunk counts
// uint16[rows * nch] | tile sums int32[tiles] | tile offsets int64[tiles+1],
// each part 256-byte aligned.
struct McScratchLayout {
    int gpl = 0, groups = 0, tiles = 0, nch = 0;
    size_t off_sum = 0, off_off = 0, need = 0;
    explicit McScratchLayout(const int32_t grid[3]) {
        if (grid[0] < 2 || grid[1] < 2 || grid[2] < 2) return;
        const int cyr = grid[1] - 1;
        gpl = (cyr + kGroupRows - 1) / kGroupRows;
        groups = gpl * (grid[2] - 1);
        tiles = (groups + kTileGroups - 1) / kTileGroups;
        nch = (grid[0] - 1 + 63) / 64;
        const size_t rows = (size_t)cyr * (size_t)(grid[2] - 1);
        off_sum = ((2 * rows * (size_t)nch + 255) / 256) * 256;
        off_off = off_sum + ((4 * (size_t)tiles + 255) / 256) * 256;
        need = off_off + 8 * ((size_t)tiles + 1);
    }
};

size_t mc_scratch_need(const int32_t grid[3]) { return McScratchLayout(grid).need; }

int mc_run(const float* field, const int32_t grid[3], const float vs[3], const float off[3],
           float* verts, int32_t* tris, int64_t capacity, int64_t* total, bool count_only,
           void** scratch, size_t* scratch_bytes, hipStream_t s, bool may_allocate) {
    *total = 0;
    if (grid[0] < 2 || grid[1] < 2 || grid[2] < 2) return RECON_OK;
    const McParams p = make_params(field, grid, vs, off);
    const McScratchLayout L(grid);
    if ((int64_t)(grid[1] - 1) * (grid[2] - 1) * L.nch >= (int64_t(1) << 31))
        return fail(RECON_ERR_RANGE, "marching cubes: more than 2^31 64-cube chunks");
    if (*scratch_bytes < L.need && !may_allocate)
        return fail(RECON_ERR_RANGE, "marching cubes needs %zu bytes of scratch, %zu available",
                    L.need, *scratch_bytes);
    if (*scratch_bytes < L.need) {
        if (*scratch) (void)hipFree(*scratch);
        *scratch = nullptr;
        *scratch_bytes = 0;
        RECON_HIP_TRY(hipMalloc(scratch, L.need), "marching cubes scratch");
        *scratch_bytes = L.need;
    }
    char* sb = static_cast<char*>(*scratch);
    uint16_t* cc = reinterpret_cast<uint16_t*>(sb);
    int32_t* tile_sum = reinterpret_cast<int32_t*>(sb + L.off_sum);
    long long* tile_off = reinterpret_cast<long long*>(sb + L.off_off);
    hipLaunchKernelGGL(mc_chunk_count_kernel, dim3(L.tiles), dim3(256), 0, s, p, L.gpl, L.groups,
                       L.nch, cc, tile_sum);
    RECON_HIP_TRY(hipGetLastError(), "mc_chunk_count_kernel launch");
    hipLaunchKernelGGL(mc_tile_scan_kernel, dim3(1), dim3(1024), 0, s, tile_sum, L.tiles, tile_off);
    RECON_HIP_TRY(hipGetLastError(), "mc_tile_scan_kernel launch");
    long long host_total = 0;
    RECON_HIP_TRY(hipMemcpyAsync(&host_total, tile_off + L.tiles, sizeof(long long),
                                 hipMemcpyDeviceToHost, s),
                  "marching cubes total");
    if (!count_only && capacity > 0 && verts) {
        hipLaunchKernelGGL(mc_chunk_emit_kernel, dim3(L.tiles), dim3(256), 0, s, p, L.gpl,
                           L.groups, L.nch, cc, tile_off, verts, tris, (long long)capacity);
        RECON_HIP_TRY(hipGetLastError(), "mc_chunk_emit_kernel launch");
    }
    RECON_HIP_TRY(hipStreamSynchronize(s), "marching cubes synchronize");
    *total = host_total;
    return RECON_OK;
}

}  // namespace recon

using namespace recon;

static int mc_entry(const float* field, const int32_t grid[3], const float vs[3], const float off[3],
                    float* verts, int32_t* tris, int64_t capacity, int64_t* n, bool count_only,
                    void** scratch, size_t* scratch_bytes, hipStream_t s, const char* fn,
                    bool may_allocate = true) {
    if (!n) return fail(RECON_ERR_NULL, "%s: num_triangles is NULL", fn);
    *n = 0;
    int rc = check_grid(field, grid, vs, off, fn);
    if (rc) return rc;
    if (!count_only) {
        if (capacity < 0) return fail(RECON_ERR_SIZE, "%s: negative capacity", fn);
        if (capacity > 0 && !verts) return fail(RECON_ERR_NULL, "%s: verts is NULL", fn);
        if (tris && capacity > (int64_t)715827882)
            return fail(RECON_ERR_RANGE, "%s: int32 vertex indices overflow above 715827882 triangles", fn);
    }
    rc = require_gfx950();
    if (rc) return rc;
    if (!may_allocate && mc_scratch_need(grid) > *scratch_bytes)
        return fail(RECON_ERR_RANGE,
                    "%s: a %d x %d x %d grid needs %zu bytes of scratch, above the static %zu",
                    fn, grid[0], grid[1], grid[2], mc_scratch_need(grid), *scratch_bytes);
    if (!may_allocate && !*scratch) return fail(RECON_ERR_HIP, "%s: no scratch address", fn);
    rc = mc_run(field, grid, vs, off, verts, tris, capacity, n, count_only, scratch, scratch_bytes,
                s, may_allocate);
    if (rc) return rc;
    if (!count_only && *n > capacity)
        return fail(RECON_ERR_CAPACITY, "%s: %lld triangles do not fit capacity %lld", fn,
                    (long long)*n, (long long)capacity);
    return RECON_OK;
}

extern "C" int recon_tsdf_count_triangles(recon_tsdf* h, int64_t* num_triangles,
                                          recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_count_triangles: handle is NULL");
    return mc_entry(h->D, h->info.grid, h->info.voxel_size, h->info.global_offset, nullptr, nullptr,
                    0, num_triangles, true, &h->mc_scratch, &h->mc_scratch_bytes,
                    (hipStream_t)stream, "recon_tsdf_count_triangles");
}

extern "C" int recon_tsdf_extract_mesh(recon_tsdf* h, float* verts, int32_t* tris,
                                       int64_t capacity_triangles, int64_t* num_triangles,
                                       recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_extract_mesh: handle is NULL");
    return mc_entry(h->D, h->info.grid, h->info.voxel_size, h->info.global_offset, verts, tris,
                    capacity_triangles, num_triangles, false, &h->mc_scratch, &h->mc_scratch_bytes,
                    (hipStream_t)stream, "recon_tsdf_extract_mesh");
}

extern "C" int recon_mc_count(const float* field, const int32_t grid[3], const float voxel_size[3],
                              const float offset[3], int64_t* num_triangles,
                              recon_stream_t stream) {
    clear_error();
    void* scratch = nullptr;
    size_t bytes = 0;
    const int rc = mc_entry(field, grid, voxel_size, offset, nullptr, nullptr, 0, num_triangles, true,
                            &scratch, &bytes, (hipStream_t)stream, "recon_mc_count");
    if (scratch) (void)hipFree(scratch);
    return rc;
}

extern "C" int recon_mc_extract(const float* field, const int32_t grid[3],
                                const float voxel_size[3], const float offset[3], float* verts,
                                int32_t* tris, int64_t capacity_triangles,
                                int64_t* num_triangles, recon_stream_t stream) {
    clear_error();
    void* scratch = nullptr;
    size_t bytes = 0;
    const int rc = mc_entry(field, grid, voxel_size, offset, verts, tris, capacity_triangles,
                            num_triangles, false, &scratch, &bytes, (hipStream_t)stream,
                            "recon_mc_extract");
    if (scratch) (void)hipFree(scratch);
    return rc;
}

extern "C" int recon_mc_class_histogram(const float* field, const int32_t grid[3],
                                        unsigned long long* hist_device, recon_stream_t stream) {
    clear_error();
    const float one[3] = {1.f, 1.f, 1.f}, zero[3] = {0.f, 0.f, 0.f};
    int rc = check_grid(field, grid, one, zero, "recon_mc_class_histogram");
    if (rc) return rc;
    if (!hist_device) return fail(RECON_ERR_NULL, "recon_mc_class_histogram: hist is NULL");
    rc = require_gfx950();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    RECON_HIP_TRY(hipMemsetAsync(hist_device, 0, 256 * sizeof(unsigned long long), s),
                  "histogram clear");
    if (grid[0] < 2 || grid[1] < 2 || grid[2] < 2) return RECON_OK;
    const McParams p = make_params(field, grid, one, zero);
    const int blocks = (p.rows + kRowsPerBlock - 1) / kRowsPerBlock;
    hipLaunchKernelGGL(mc_class_hist_kernel, dim3(blocks), dim3(256), 0, s, p, hist_device);
    RECON_HIP_TRY(hipGetLastError(), "mc_class_hist_kernel launch");
    return RECON_OK;
}
