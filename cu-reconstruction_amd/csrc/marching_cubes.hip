// marching_cubes.hip -- marching cubes over a dense TSDF for gfx950.
//
// Replaces generate_triangular_surface (cu-Reconstruction
// modules/src/cuda/marching_cubes.cu:102-236).  The reference launches one
// kernel per z layer, synchronises, and compacts each layer on the host from
// managed memory.  Here the whole volume is done on the device in four
// launches with a fixed, deterministic output order:
//   1. chunk count - one wavefront per 4 cube rows (y..y+3, z), streaming
//                    their 10 voxel rows once, 256 cubes a step (16-byte
//                    loads; 64 with 4-byte loads where rows are not 16-byte
//                    aligned): every
//                    cube's class (marching_cubes.cu:76-84,118-130) from
//                    the corners' sign bits, its triangle count from a
//                    table, and per 64-cube chunk of each row the chunk's
//                    triangle count; per workgroup ("tile", 16 cube rows)
//                    the tile's total and its number of chunks with
//                    triangles;
//   2. tile scan   - one workgroup: exclusive scans of the tile totals
//                    (each tile's first triangle and the volume's total)
//                    and of the tiles' groups of four chunks with
//                    triangles;
//   3. group prep  - one wavefront per tile: a 16-byte record per group of
//                    four chunks with triangles (its tile, first triangle
//                    and chunks);
//   4. emit        - a grid sized to the resident workgroups: each
//                    wavefront takes groups gi = wave, wave + waves, ...,
//                    reads each group's record and spreads its triangles
//                    (16 lanes per chunk, 4 cubes per lane) over its 64
//                    lanes to compute the vertices (marching_cubes.cu:23-74),
//                    staged in LDS and written in consecutive dwords.  A tile
//                    dense with triangles (a floor along a cube layer) is
//                    thus shared by many wavefronts.  The allocation-free
//                    entry skips step 3 (its static scratch has no room for
//                    the records) and finds each group's tile and chunks by
//                    a search and a walk instead.
// Output order is cube z, y, x, table slot -- the order the reference's host
// loop produces (marching_cubes.cu:163-188 inside :213-225) -- as a triangle
// soup: triangle t owns vertices 3t, 3t+1, 3t+2 (marching_cubes.cu:171).
#pragma clang fp contract(off)

#include "mc_device.hpp"

#include <algorithm>
#include <mutex>

namespace recon {

constexpr int kRowsPerBlock = 4;  // one wavefront per row, 256 threads

// ---- tile scan ------------------------------------------------------------------
// The scan takes kScanRows rows of kScanRow tiles a pass, thread t holding
// tiles 4t .. 4t+3 of each row, so every load and store instruction of a
// wavefront covers contiguous bytes (one workgroup's loads and stores all go
// through one CU).
constexpr int kScanRow = 4096, kScanRows = 2;

// One workgroup of 1024 threads: tile_off[i] = sum of tile_sum[0..i) and
// tile_off[n] = the total; when tile_nne is not null also group_off[i] = sum
// over tiles before i of ceil(tile_nne / 4) (the emit's four-chunk groups)
// and group_off[n] = their total.  All four arrays are 16-byte aligned.
__global__ __launch_bounds__(1024) void mc_tile_scan_kernel(const int32_t* __restrict__ tile_sum,
                                                            const int32_t* __restrict__ tile_nne,
                                                            int n, long long* __restrict__ tile_off,
                                                            int32_t* __restrict__ group_off) {
    __shared__ long long wtot[kScanRows][16];
    __shared__ int gtot[kScanRows][16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    long long carry = 0;
    int gcarry = 0;
    for (int base = 0; base < n; base += kScanRow * kScanRows) {
        int v[kScanRows][4], g[kScanRows][4];
        long long sum[kScanRows], incl[kScanRows];
        int gsum[kScanRows], ginc[kScanRows];
#pragma unroll
        for (int q = 0; q < kScanRows; ++q) {
            const int i0 = base + kScanRow * q + 4 * t;
            if (i0 + 4 <= n) {
                const int4 a = *reinterpret_cast<const int4*>(tile_sum + i0);
                v[q][0] = a.x; v[q][1] = a.y; v[q][2] = a.z; v[q][3] = a.w;
                if (tile_nne) {
                    const int4 b = *reinterpret_cast<const int4*>(tile_nne + i0);
                    g[q][0] = (b.x + 3) >> 2; g[q][1] = (b.y + 3) >> 2;
                    g[q][2] = (b.z + 3) >> 2; g[q][3] = (b.w + 3) >> 2;
                } else {
                    g[q][0] = g[q][1] = g[q][2] = g[q][3] = 0;
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    v[q][k] = i0 + k < n ? tile_sum[i0 + k] : 0;
                    g[q][k] = (tile_nne && i0 + k < n) ? (tile_nne[i0 + k] + 3) >> 2 : 0;
                }
            }
            sum[q] = (long long)v[q][0] + v[q][1] + v[q][2] + v[q][3];
            gsum[q] = g[q][0] + g[q][1] + g[q][2] + g[q][3];
        }
#pragma unroll
        for (int q = 0; q < kScanRows; ++q) {
            incl[q] = wave_inclusive_scan64(sum[q]);
            ginc[q] = wave_inclusive_scan(gsum[q], lane);
        }
        if (lane == 63) {
#pragma unroll
            for (int q = 0; q < kScanRows; ++q) {
                wtot[q][wave] = incl[q];
                gtot[q][wave] = ginc[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kScanRows; ++q) {
            long long before = 0, all = 0;
            int gbefore = 0, gall = 0;
#pragma unroll
            for (int w = 0; w < 16; ++w) {
                before += w < wave ? wtot[q][w] : 0;
                all += wtot[q][w];
                gbefore += w < wave ? gtot[q][w] : 0;
                gall += gtot[q][w];
            }
            const long long e0 = carry + before + incl[q] - sum[q];
            const long long e1 = e0 + v[q][0], e2 = e1 + v[q][1], e3 = e2 + v[q][2];
            const int f0 = gcarry + gbefore + ginc[q] - gsum[q];
            const int f1 = f0 + g[q][0], f2 = f1 + g[q][1], f3 = f2 + g[q][2];
            const int i0 = base + kScanRow * q + 4 * t;
            if (i0 + 4 <= n) {
                longlong2* d2 = reinterpret_cast<longlong2*>(tile_off + i0);
                d2[0] = make_longlong2(e0, e1);
                d2[1] = make_longlong2(e2, e3);
                if (group_off) *reinterpret_cast<int4*>(group_off + i0) = make_int4(f0, f1, f2, f3);
            } else {
                const long long e[4] = {e0, e1, e2, e3};
                const int f[4] = {f0, f1, f2, f3};
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (i0 + k < n) {
                        tile_off[i0 + k] = e[k];
                        if (group_off) group_off[i0 + k] = f[k];
                    }
            }
            carry += all;
            gcarry += gall;
        }
        __syncthreads();  // wtot and gtot are rewritten by the next pass
    }
    if (t == 0) {
        tile_off[n] = carry;
        if (group_off) group_off[n] = gcarry;
    }
}

int launch_tile_scan(const int32_t* tile_sum, int n, long long* tile_off, hipStream_t s,
                     const int32_t* tile_nne, int32_t* group_off) {
    hipLaunchKernelGGL(mc_tile_scan_kernel, dim3(1), dim3(1024), 0, s, tile_sum, tile_nne, n,
                       tile_off, group_off);
    RECON_HIP_TRY(hipGetLastError(), "mc_tile_scan_kernel launch");
    return RECON_OK;
}

// ---- single-volume kernels ------------------------------------------------------
__global__ __launch_bounds__(256) void mc_chunk_count_kernel(const McParams p, const McTiling tl,
                                                             uint16_t* __restrict__ cc,
                                                             int32_t* __restrict__ tile_sum,
                                                             int32_t* __restrict__ tile_nne) {
    __shared__ uint8_t ntab[256];
    __shared__ int wtot[8];
    ntab[threadIdx.x] = mc_ntab_qbyte((int)threadIdx.x);
    __syncthreads();
    mc_count_tile(p, tl, blockIdx.x, cc, ntab, wtot, tile_sum + blockIdx.x, tile_nne + blockIdx.x);
}

// The same with 16-byte loads (mc_count_tile4): gx % 4 == 0 and a 16-byte
// aligned field.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void mc_chunk_count4_kernel(const McParams p, const McTiling tl,
                                                              uint16_t* __restrict__ cc,
                                                              int32_t* __restrict__ tile_sum,
                                                              int32_t* __restrict__ tile_nne) {
    __shared__ uint8_t ntab[256];
    __shared__ int wtot[8];
    ntab[threadIdx.x] = mc_ntab_qbyte((int)threadIdx.x);
    __syncthreads();
    mc_count_tile4(p, tl, blockIdx.x, cc, ntab, wtot, tile_sum + blockIdx.x, tile_nne + blockIdx.x);
}

// One wavefront per tile: the records of the tile's groups (mc_group_prep).
__global__ __launch_bounds__(256) void mc_group_prep_kernel(const McParams p, const McTiling tl,
                                                            const uint16_t* __restrict__ cc,
                                                            const long long* __restrict__ tile_off,
                                                            const int32_t* __restrict__ tile_nne,
                                                            const int32_t* __restrict__ group_off,
                                                            McGroupRec* __restrict__ rec) {
    const int b = (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (b < tl.tiles)
        mc_group_prep(p, tl, b, cc, tile_off, tile_nne, group_off, rec, (int)(threadIdx.x & 63));
}

// A grid sized to the resident workgroups walks the volume's groups of four
// chunks with triangles, wavefront by wavefront (gi = wave, wave + waves,
// ...), each from its record (one scalar load, which unlike a vector load
// does not wait for the wavefront's earlier stores): the groups of a tile
// dense with triangles (a floor along a cube layer) go to as many
// wavefronts.
__global__ __launch_bounds__(256) void mc_chunk_emit_kernel(const McParams p, const McTiling tl,
                                                            const McGroupRec* __restrict__ rec,
                                                            const int32_t* __restrict__ group_off,
                                                            float* __restrict__ verts,
                                                            int32_t* __restrict__ tris,
                                                            long long capacity) {
    __shared__ McEmitLds lds;
    mc_emit_lds_init(lds);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int ngroups = group_off[tl.tiles];
    const int waves = (int)gridDim.x * 4;
    for (int gi = (int)blockIdx.x * 4 + wave; gi < ngroups; gi += waves)
        mc_emit_group_rec(p, tl, rec[gi], lds, wave, lane, verts, tris, capacity);
}

// The same walk for the allocation-free entry (no room for records): each
// wavefront finds its group's tile by a search over group_off and its
// chunks by a walk over the tile's chunk counts (mc_emit_group_at).
__global__ __launch_bounds__(256) void mc_chunk_emit_search_kernel(const McParams p, const McTiling tl,
                                                            const uint16_t* __restrict__ cc,
                                                            const long long* __restrict__ tile_off,
                                                            const int32_t* __restrict__ tile_nne,
                                                            const int32_t* __restrict__ group_off,
                                                            float* __restrict__ verts,
                                                            int32_t* __restrict__ tris,
                                                            long long capacity) {
    __shared__ McEmitLds lds;
    mc_emit_lds_init(lds);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ngroups = group_off[tl.tiles];
    const int waves = (int)gridDim.x * 4;
    for (int gi = (int)blockIdx.x * 4 + wave; gi < ngroups; gi += waves)
        mc_emit_group_at(p, tl, gi, cc, tile_off, tile_nne, group_off, lds, wave, lane, verts,
                         tris, capacity);
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

__global__ __launch_bounds__(256) void scan_block_sums_kernel(const ScanJobs j, int n) {
    __shared__ long long sh[256];
    const int32_t* __restrict__ in = blockIdx.y ? j.in[1] : j.in[0];
    const int b0 = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    long long s = 0;
    if (b0 + kScanItems <= n && (reinterpret_cast<uintptr_t>(in) & 15) == 0) {
        const int4 a = *reinterpret_cast<const int4*>(in + b0);
        s = ((long long)a.x + a.y) + ((long long)a.z + a.w);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (b0 + k < n) s += in[b0 + k];
    }
    long long total;
    block_exclusive_scan(s, sh, &total);
    if (threadIdx.x == 0) (blockIdx.y ? j.sums[1] : j.sums[0])[blockIdx.x] = total;
}

// One workgroup per array: exclusive scan of its block sums in place, total
// at [nb]; each thread takes kTopItems consecutive sums per pass, so up to
// 256 * kTopItems sums need one workgroup scan.
constexpr int kTopItems = 8;
__global__ __launch_bounds__(256) void scan_top_kernel(const ScanJobs j, int nb) {
    __shared__ long long sh[256];
    long long* __restrict__ sums = blockIdx.y ? j.sums[1] : j.sums[0];
    long long carry = 0;
    for (int base = 0; base < nb; base += 256 * kTopItems) {
        const int i0 = base + threadIdx.x * kTopItems;
        long long v[kTopItems], s = 0;
#pragma unroll
        for (int k = 0; k < kTopItems; ++k) {
            v[k] = i0 + k < nb ? sums[i0 + k] : 0;
            s += v[k];
        }
        long long total;
        long long e = carry + block_exclusive_scan(s, sh, &total);
#pragma unroll
        for (int k = 0; k < kTopItems; ++k) {
            if (i0 + k < nb) sums[i0 + k] = e;
            e += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ __launch_bounds__(256) void scan_apply_kernel(const ScanJobs j, int n) {
    __shared__ long long sh[256];
    const int32_t* __restrict__ in = blockIdx.y ? j.in[1] : j.in[0];
    long long* __restrict__ out = blockIdx.y ? j.out[1] : j.out[0];
    const int b0 = blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    const bool vec = b0 + kScanItems <= n && (reinterpret_cast<uintptr_t>(in) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    long long vals[kScanItems];
    if (vec) {
        const int4 a = *reinterpret_cast<const int4*>(in + b0);
        vals[0] = a.x; vals[1] = a.y; vals[2] = a.z; vals[3] = a.w;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) vals[k] = (b0 + k < n) ? in[b0 + k] : 0;
    }
    const long long s = (vals[0] + vals[1]) + (vals[2] + vals[3]);
    long long total;
    const long long e0 =
        block_exclusive_scan(s, sh, &total) + (blockIdx.y ? j.sums[1] : j.sums[0])[blockIdx.x];
    const long long e1 = e0 + vals[0], e2 = e1 + vals[1], e3 = e2 + vals[2];
    if (vec) {
        longlong2* o2 = reinterpret_cast<longlong2*>(out + b0);
        o2[0] = make_longlong2(e0, e1);
        o2[1] = make_longlong2(e2, e3);
    } else {
        const long long e[kScanItems] = {e0, e1, e2, e3};
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (b0 + k < n) out[b0 + k] = e[k];
    }
}

int launch_exclusive_scans(const ScanJobs& j, int count, int n, hipStream_t s) {
    if (count < 1 || count > 2) return fail(RECON_ERR_SIZE, "launch_exclusive_scans: %d arrays", count);
    const int nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(scan_block_sums_kernel, dim3(nb, count), dim3(256), 0, s, j, n);
    RECON_HIP_TRY(hipGetLastError(), "scan_block_sums_kernel launch");
    hipLaunchKernelGGL(scan_top_kernel, dim3(1, count), dim3(256), 0, s, j, nb);
    RECON_HIP_TRY(hipGetLastError(), "scan_top_kernel launch");
    bool apply = true;
    for (int i = 0; i < count; ++i) apply = apply && j.out[i];
    if (apply) {
        hipLaunchKernelGGL(scan_apply_kernel, dim3(nb, count), dim3(256), 0, s, j, n);
        RECON_HIP_TRY(hipGetLastError(), "scan_apply_kernel launch");
    }
    return RECON_OK;
}

int launch_exclusive_scan(const int32_t* counts, int n, long long* offsets, long long* sums,
                          hipStream_t s) {
    ScanJobs j;
    j.in[0] = j.in[1] = counts;
    j.sums[0] = j.sums[1] = sums;
    j.out[0] = j.out[1] = offsets;
    return launch_exclusive_scans(j, 1, n, s);
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
// uint16[rows * nch] | tile sums int32[tiles] | tile chunks with triangles
// int32[tiles] | tile offsets int64[tiles+1] | tile first groups
// int32[tiles+1] (need bytes, the search emit's) | group records
// McGroupRec[4 * nch per tile] (need_rec bytes in all, the record emit's),
// each part 256-byte aligned.  Records fit their fields when the tile index
// stays below 2^24, the triangle count below 2^40 (at most 5 per cube) and a
// tile's chunks below kNoChunk.
struct McScratchLayout {
    McTiling t{0, 0, 0, 0};
    int gpl = 0, groups = 0, tiles = 0, nch = 0;
    size_t off_sum = 0, off_nne = 0, off_off = 0, off_goff = 0, need = 0;
    size_t off_rec = 0, need_rec = 0;
    bool rec_ok = false;
    explicit McScratchLayout(const int32_t grid[3]) {
        if (grid[0] < 2 || grid[1] < 2 || grid[2] < 2) return;
        t = mc_tiling(grid[0], grid[1], grid[2]);
        gpl = t.gpl;
        groups = t.groups;
        tiles = t.tiles;
        nch = t.nch;
        const size_t rows = (size_t)(grid[1] - 1) * (size_t)(grid[2] - 1);
        auto a256 = [](size_t b) { return (b + 255) / 256 * 256; };
        off_sum = a256(2 * rows * (size_t)nch);
        off_nne = off_sum + a256(4 * (size_t)tiles);
        off_off = off_nne + a256(4 * (size_t)tiles);
        off_goff = off_off + a256(8 * ((size_t)tiles + 1));
        need = off_goff + 4 * ((size_t)tiles + 1);
        off_rec = a256(need);
        need_rec = off_rec + sizeof(McGroupRec) * (size_t)tiles * 4 * (size_t)nch;
        const double cubes = (double)(grid[0] - 1) * (double)rows;
        rec_ok = tiles < (1 << 24) && 5.0 * cubes < 0x1p40 && 16 * (size_t)nch < kNoChunk;
    }
};

// Workgroups of an emit pass: its resident slots on every CU of the current
// device (looked up once per device and kernel; which = 0 the record emit,
// 1 the search emit).
static int emit_grid(int which) {
    static std::mutex mu;
    static int grid[2][64] = {{0}};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
    std::lock_guard<std::mutex> lk(mu);
    if (!grid[which][dev]) {
        int cus = 0, occ = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus < 1)
            cus = 256;
        const hipError_t e =
            which == 0
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, mc_chunk_emit_kernel, 256, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, mc_chunk_emit_search_kernel,
                                                               256, 0);
        if (e != hipSuccess || occ < 1) occ = 1;
        grid[which][dev] = cus * occ;
    }
    return grid[which][dev];
}

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
    // the record emit when allocating; the allocation-free entry keeps to the
    // search emit's smaller scratch
    const bool use_rec = L.rec_ok && may_allocate;
    const size_t need = use_rec ? L.need_rec : L.need;
    if (*scratch_bytes < need) {
        if (*scratch) (void)hipFree(*scratch);
        *scratch = nullptr;
        *scratch_bytes = 0;
        RECON_HIP_TRY(hipMalloc(scratch, need), "marching cubes scratch");
        *scratch_bytes = need;
    }
    char* sb = static_cast<char*>(*scratch);
    uint16_t* cc = reinterpret_cast<uint16_t*>(sb);
    int32_t* tile_sum = reinterpret_cast<int32_t*>(sb + L.off_sum);
    int32_t* tile_nne = reinterpret_cast<int32_t*>(sb + L.off_nne);
    long long* tile_off = reinterpret_cast<long long*>(sb + L.off_off);
    int32_t* group_off = reinterpret_cast<int32_t*>(sb + L.off_goff);
    McGroupRec* rec = reinterpret_cast<McGroupRec*>(sb + L.off_rec);
    if (grid[0] % 4 == 0 && reinterpret_cast<uintptr_t>(field) % 16 == 0)
        hipLaunchKernelGGL(mc_chunk_count4_kernel, dim3(L.tiles), dim3(256), 0, s, p, L.t, cc,
                           tile_sum, tile_nne);
    else
        hipLaunchKernelGGL(mc_chunk_count_kernel, dim3(L.tiles), dim3(256), 0, s, p, L.t, cc,
                           tile_sum, tile_nne);
    RECON_HIP_TRY(hipGetLastError(), "mc_chunk_count_kernel launch");
    int rc = launch_tile_scan(tile_sum, L.tiles, tile_off, s, tile_nne, group_off);
    if (rc) return rc;
    // The emit checks the capacity on the device, so every launch is queued
    // before the host waits: the copy of the total to pageable memory comes
    // last, since it returns only when the stream has reached it.
    if (!count_only && capacity > 0 && verts) {
        if (use_rec) {
            hipLaunchKernelGGL(mc_group_prep_kernel, dim3((L.tiles + 3) / 4), dim3(256), 0, s, p,
                               L.t, cc, tile_off, tile_nne, group_off, rec);
            RECON_HIP_TRY(hipGetLastError(), "mc_group_prep_kernel launch");
            hipLaunchKernelGGL(mc_chunk_emit_kernel, dim3(emit_grid(0)), dim3(256), 0, s, p, L.t,
                               rec, group_off, verts, tris, (long long)capacity);
            RECON_HIP_TRY(hipGetLastError(), "mc_chunk_emit_kernel launch");
        } else {
            hipLaunchKernelGGL(mc_chunk_emit_search_kernel, dim3(emit_grid(1)), dim3(256), 0, s, p,
                               L.t, cc, tile_off, tile_nne, group_off, verts, tris,
                               (long long)capacity);
            RECON_HIP_TRY(hipGetLastError(), "mc_chunk_emit_search_kernel launch");
        }
    }
    long long host_total = 0;
    RECON_HIP_TRY(hipMemcpyAsync(&host_total, tile_off + L.tiles, sizeof(long long),
                                 hipMemcpyDeviceToHost, s),
                  "marching cubes total");
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

// ---- marching cubes without allocation ---------------------------------------
// Scratch of recon_mc_extract_noalloc: the chunk counts (2 B per 64-cube
// chunk of every cube row) and the tile sums, chunk counts, triangle and
// group offsets (20 B per 16 cube rows) -- mc_scratch_need bytes -- in one
// static array that each GPU holds
// for the life of the process.  48 MiB covers a 1024^3 grid (1023^2 cube
// rows of 16 chunks: 34.6 MB).
constexpr size_t kMcNoAllocScratchBytes = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kMcNoAllocScratchBytes];

// The array's device address on the current device, looked up once per
// device (hipGetSymbolAddress; the host-side name is not a device address).
static int noalloc_scratch(void** out) {
    static std::mutex mu;
    static void* addr[64] = {nullptr};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= 64) return fail(RECON_ERR_DEVICE, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(marching-cubes scratch)");
        addr[dev] = p;
    }
    *out = addr[dev];
    return RECON_OK;
}

extern "C" int recon_mc_extract_noalloc(const float* field, const int32_t grid[3],
                                        const float voxel_size[3], const float offset[3],
                                        float* verts, int32_t* tris, int64_t capacity_triangles,
                                        int64_t* num_triangles, recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_mc_extract_noalloc";
    if (!num_triangles) return fail(RECON_ERR_NULL, "%s: num_triangles is NULL", fn);
    *num_triangles = 0;
    int rc = check_grid(field, grid, voxel_size, offset, fn);
    if (rc) return rc;
    if (mc_scratch_need(grid) > kMcNoAllocScratchBytes)
        return fail(RECON_ERR_RANGE,
                    "%s: a %d x %d x %d grid needs %zu bytes of scratch, above the static maximum "
                    "of %zu (2 B per 64-cube chunk of each of the (gy-1)*(gz-1) cube rows)",
                    fn, grid[0], grid[1], grid[2], mc_scratch_need(grid), kMcNoAllocScratchBytes);
    rc = require_gfx950();
    if (rc) return rc;
    void* sp = nullptr;
    rc = noalloc_scratch(&sp);
    if (rc) return rc;
    size_t bytes = kMcNoAllocScratchBytes;
    return mc_entry(field, grid, voxel_size, offset, verts, tris, capacity_triangles, num_triangles,
                    false, &sp, &bytes, (hipStream_t)stream, fn, false);
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
