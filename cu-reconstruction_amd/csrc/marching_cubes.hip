// marching_cubes.hip -- marching cubes over a dense TSDF for gfx950.
//
// Replaces generate_triangular_surface (cu-Reconstruction
// modules/src/cuda/marching_cubes.cu:102-236).  The reference launches one
// kernel per z layer, synchronises, and compacts each layer on the host from
// managed memory.  Here the whole volume is done on the device in four
// launches with a fixed, deterministic output order:
//   1. row_count  - one wavefront per 4 cube rows (y..y+3, z): classify every
//                   cube (marching_cubes.cu:76-84,118-130) and count each
//                   row's triangles;
//   2-4. scan     - exclusive scan of the row counts (block sums, top-level
//                   scan, apply) giving each row's first triangle;
//   5. emit       - rows with triangles only: a wave-level prefix sum of the
//                   per-cube counts places every triangle, then each lane
//                   writes its triangles' vertices (marching_cubes.cu:23-74).
// Output order is cube z, y, x, table slot -- the order the reference's host
// loop produces (marching_cubes.cu:163-188 inside :213-225) -- as a triangle
// soup: triangle t owns vertices 3t, 3t+1, 3t+2 (marching_cubes.cu:171).
#pragma clang fp contract(off)

#include "mc_device.hpp"

#include <algorithm>

namespace recon {

constexpr int kRowsPerBlock = 4;  // one wavefront per row, 256 threads

// Cube rows per wavefront in the count pass: 4 rows of one layer read 10
// voxel rows instead of 16 (count 207 -> 152 us at 512^3; 2 rows: 165 us;
// the chunk masks add 8 us and save the emit 27 us).
constexpr int kCountRows = 4;

// Triangles of every cube row, and a chunk mask per row for the emit pass:
// chunk_bit of the 64-cube chunks holding a triangle in the row's group of
// kCountRows rows (a superset of the row's own, one ballot per step).
__global__ __launch_bounds__(256) void mc_row_count_kernel(const McParams p,
                                                           int32_t* __restrict__ row_count,
                                                           uint32_t* __restrict__ row_mask) {
    const int lane = threadIdx.x & 63;
    const int groups = (p.cyr + kCountRows - 1) / kCountRows;
    const int q = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    const int z = q / groups, y = kCountRows * (q % groups);
    if (z >= p.rows / p.cyr) return;
    const int r = y + p.cyr * z;
    if (y + kCountRows > p.cyr) {  // the layer's last, partial group: one row at a time
        for (int i = 0; y + i < p.cyr; ++i) {
            uint32_t m;
            const int n = mc_row_triangles(p, r + i, lane, &m);
            if (lane == 0) {
                row_count[r + i] = n;
                row_mask[r + i] = m;
            }
        }
        return;
    }
    CornerStreamK<kCountRows> cs(p, y, z, lane);
    int n[kCountRows];
#pragma unroll
    for (int i = 0; i < kCountRows; ++i) n[i] = 0;
    uint32_t m = 0;  // chunks with a triangle in any of the group's rows
    for (int x0 = 0; x0 < p.cx; x0 += 64) {
        float v[kCountRows][8];
        cs.step(x0, lane, v);
        if (x0 + lane < p.cx) {
            int any = 0;
#pragma unroll
            for (int i = 0; i < kCountRows; ++i) {
                const int t = class_ntri(cube_class(v[i]));
                n[i] += t;
                any |= t;
            }
            if (__ballot(any != 0)) m |= chunk_bit(x0);
        }
    }
#pragma unroll
    for (int i = 0; i < kCountRows; ++i) {
        const int t = wave_sum(n[i]);
        if (lane == 0) {
            row_count[r + i] = t;
            row_mask[r + i] = m;
        }
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

__global__ __launch_bounds__(256) void mc_emit_kernel(const McParams p,
                                                      const int32_t* __restrict__ row_count,
                                                      const uint32_t* __restrict__ row_mask,
                                                      const long long* __restrict__ row_offset,
                                                      float* __restrict__ verts,
                                                      int32_t* __restrict__ tris,
                                                      long long capacity) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * kRowsPerBlock + (threadIdx.x >> 6);
    if (r >= p.rows) return;
    if (row_count[r] == 0) return;
    mc_emit_row_masked(p, r, lane, row_offset[r], row_mask[r], verts, tris, capacity);
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

// Scratch layout: row_count int32[rows] | row_mask uint32[rows] |
// row_offset int64[rows] | sums int64[nb+1], each part 256-byte aligned
int mc_run(const float* field, const int32_t grid[3], const float vs[3], const float off[3],
           float* verts, int32_t* tris, int64_t capacity, int64_t* total, bool count_only,
           void** scratch, size_t* scratch_bytes, hipStream_t s) {
    *total = 0;
    if (grid[0] < 2 || grid[1] < 2 || grid[2] < 2) return RECON_OK;
    const McParams p = make_params(field, grid, vs, off);
    const int rows = p.rows;
    const int nb = (rows + kScanBlock - 1) / kScanBlock;
    const size_t off_count = 0;
    const size_t off_mask = ((sizeof(int32_t) * (size_t)rows + 255) / 256) * 256;
    const size_t off_offset = off_mask + ((sizeof(uint32_t) * (size_t)rows + 255) / 256) * 256;
    const size_t off_sums = off_offset + ((sizeof(long long) * (size_t)rows + 255) / 256) * 256;
    const size_t need = off_sums + sizeof(long long) * (size_t)(nb + 1);
    if (*scratch_bytes < need) {
        if (*scratch) (void)hipFree(*scratch);
        *scratch = nullptr;
        *scratch_bytes = 0;
        RECON_HIP_TRY(hipMalloc(scratch, need), "marching cubes scratch");
        *scratch_bytes = need;
    }
    char* sb = static_cast<char*>(*scratch);
    int32_t* row_count = reinterpret_cast<int32_t*>(sb + off_count);
    uint32_t* row_mask = reinterpret_cast<uint32_t*>(sb + off_mask);
    long long* row_offset = reinterpret_cast<long long*>(sb + off_offset);
    long long* sums = reinterpret_cast<long long*>(sb + off_sums);
    const int blocks = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
    const int count_blocks =
        (((p.cyr + kCountRows - 1) / kCountRows) * (grid[2] - 1) + kRowsPerBlock - 1) / kRowsPerBlock;
    hipLaunchKernelGGL(mc_row_count_kernel, dim3(count_blocks), dim3(256), 0, s, p, row_count,
                       row_mask);
    RECON_HIP_TRY(hipGetLastError(), "mc_row_count_kernel launch");
    int rc = launch_exclusive_scan(row_count, rows, count_only ? nullptr : row_offset, sums, s);
    if (rc) return rc;
    long long host_total = 0;
    RECON_HIP_TRY(hipMemcpyAsync(&host_total, sums + nb, sizeof(long long), hipMemcpyDeviceToHost, s),
                  "marching cubes total");
    if (!count_only) {
        if (capacity > 0 && verts) {
            hipLaunchKernelGGL(mc_emit_kernel, dim3(blocks), dim3(256), 0, s, p, row_count,
                               row_mask, row_offset, verts, tris, (long long)capacity);
            RECON_HIP_TRY(hipGetLastError(), "mc_emit_kernel launch");
        }
    }
    RECON_HIP_TRY(hipStreamSynchronize(s), "marching cubes synchronize");
    *total = host_total;
    return RECON_OK;
}

}  // namespace recon

using namespace recon;

static int mc_entry(const float* field, const int32_t grid[3], const float vs[3], const float off[3],
                    float* verts, int32_t* tris, int64_t capacity, int64_t* n, bool count_only,
                    void** scratch, size_t* scratch_bytes, hipStream_t s, const char* fn) {
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
    rc = mc_run(field, grid, vs, off, verts, tris, capacity, n, count_only, scratch, scratch_bytes,
                s);
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
