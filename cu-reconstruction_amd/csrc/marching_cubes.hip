// marching_cubes.hip -- marching cubes over a dense TSDF for gfx950.
//
// Replaces generate_triangular_surface (cu-Reconstruction
// modules/src/cuda/marching_cubes.cu:102-236).  The reference launches one
// kernel per z layer, synchronises, and compacts each layer on the host from
// managed memory.  Here the whole volume is done on the device in four
// launches with a fixed, deterministic output order:
//   1. row_count  - one wavefront per cube row (y,z): classify every cube
//                   (marching_cubes.cu:76-84,118-130) and count its triangles;
//   2-4. scan     - exclusive scan of the row counts (block sums, top-level
//                   scan, apply) giving each row's first triangle;
//   5. emit       - rows with triangles only: a wave-level prefix sum of the
//                   per-cube counts places every triangle, then each lane
//                   writes its triangles' vertices (marching_cubes.cu:23-74).
// Output order is cube z, y, x, table slot -- the order the reference's host
// loop produces (marching_cubes.cu:163-188 inside :213-225) -- as a triangle
// soup: triangle t owns vertices 3t, 3t+1, 3t+2 (marching_cubes.cu:171).
#pragma clang fp contract(off)

#include "recon_common.hpp"

#include "recon/mc_tables.h"

#include <algorithm>

namespace recon {

__constant__ uint64_t c_tri[256] = RECON_MC_TRI_PACKED_INIT;
__constant__ int c_edge[12][2] = RECON_MC_EDGE_CORNERS_INIT;

struct McParams {
    const float* f;
    int32_t gx, gy, gz;
    int32_t cx;        // cubes per row = gx - 1
    int32_t cyr;       // cube rows per layer = gy - 1
    int32_t rows;      // (gy - 1) * (gz - 1)
    float vs[3];
    float off[3];
};

constexpr int kRowsPerBlock = 4;  // one wavefront per row, 256 threads

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

__device__ __forceinline__ int row_of_block(int lane_row) {
    return blockIdx.x * kRowsPerBlock + lane_row;
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

__global__ __launch_bounds__(256) void mc_row_count_kernel(const McParams p,
                                                           int32_t* __restrict__ row_count) {
    const int lane = threadIdx.x & 63;
    const int r = row_of_block(threadIdx.x >> 6);
    if (r >= p.rows) return;
    const int y = r % p.cyr, z = r / p.cyr;
    int n = 0;
    for (int x = lane; x < p.cx; x += 64) {
        float v[8];
        load_corners(p, x, y, z, v);
        n += class_ntri(cube_class(v));
    }
    n = wave_sum(n);
    if (lane == 0) row_count[r] = n;
}

__global__ __launch_bounds__(256) void mc_class_hist_kernel(const McParams p,
                                                            unsigned long long* __restrict__ hist) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int r = row_of_block(threadIdx.x >> 6);
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

// ---- exclusive scan of int32 row counts into int64 offsets -----------------
constexpr int kScanItems = 4;
constexpr int kScanBlock = 256 * kScanItems;

__device__ __forceinline__ long long block_exclusive_scan(long long v, long long* sh,
                                                          long long* total) {
    // sh: 256 entries of shared scratch
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const long long o = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += o;
        __syncthreads();
    }
    const long long incl = sh[t];
    *total = sh[255];
    __syncthreads();
    return incl - v;
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
    // corner offsets packed as bits: dx = bit0, dy = bit1, dz = bit2 of k_off
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

__global__ __launch_bounds__(256) void mc_emit_kernel(const McParams p,
                                                      const int32_t* __restrict__ row_count,
                                                      const long long* __restrict__ row_offset,
                                                      float* __restrict__ verts,
                                                      int32_t* __restrict__ tris,
                                                      long long capacity) {
    const int lane = threadIdx.x & 63;
    const int r = row_of_block(threadIdx.x >> 6);
    if (r >= p.rows) return;
    if (row_count[r] == 0) return;
    const int y = r % p.cyr, z = r / p.cyr;
    long long base = row_offset[r];
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
                    if (tris) tris[3 * t + k] = (int32_t)(3 * t + k);
                }
            }
        }
        base += chunk_total;
    }
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

// Scratch layout: row_count int32[rows] | row_offset int64[rows] | sums int64[nb+1]
int mc_run(const float* field, const int32_t grid[3], const float vs[3], const float off[3],
           float* verts, int32_t* tris, int64_t capacity, int64_t* total, bool count_only,
           void** scratch, size_t* scratch_bytes, hipStream_t s) {
    *total = 0;
    if (grid[0] < 2 || grid[1] < 2 || grid[2] < 2) return RECON_OK;
    const McParams p = make_params(field, grid, vs, off);
    const int rows = p.rows;
    const int nb = (rows + kScanBlock - 1) / kScanBlock;
    const size_t off_count = 0;
    const size_t off_offset = ((sizeof(int32_t) * (size_t)rows + 255) / 256) * 256;
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
    long long* row_offset = reinterpret_cast<long long*>(sb + off_offset);
    long long* sums = reinterpret_cast<long long*>(sb + off_sums);
    const int blocks = (rows + kRowsPerBlock - 1) / kRowsPerBlock;
    hipLaunchKernelGGL(mc_row_count_kernel, dim3(blocks), dim3(256), 0, s, p, row_count);
    RECON_HIP_TRY(hipGetLastError(), "mc_row_count_kernel launch");
    hipLaunchKernelGGL(scan_block_sums_kernel, dim3(nb), dim3(256), 0, s, row_count, rows, sums);
    RECON_HIP_TRY(hipGetLastError(), "scan_block_sums_kernel launch");
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(256), 0, s, sums, nb);
    RECON_HIP_TRY(hipGetLastError(), "scan_top_kernel launch");
    long long host_total = 0;
    RECON_HIP_TRY(hipMemcpyAsync(&host_total, sums + nb, sizeof(long long), hipMemcpyDeviceToHost, s),
                  "marching cubes total");
    if (!count_only) {
        hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(256), 0, s, row_count, rows, sums,
                           row_offset);
        RECON_HIP_TRY(hipGetLastError(), "scan_apply_kernel launch");
        if (capacity > 0 && verts) {
            hipLaunchKernelGGL(mc_emit_kernel, dim3(blocks), dim3(256), 0, s, p, row_count,
                               row_offset, verts, tris, (long long)capacity);
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
