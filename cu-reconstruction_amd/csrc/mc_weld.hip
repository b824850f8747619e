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
// which the vertex writer and the triangle writer both evaluate: 28 B of
// masks and counts and 8 B of base per 64 voxels.
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

constexpr int kWeldWaves = 4;  // one wavefront per chunk or cube row, 256 threads

struct WeldParams {
    McParams mc;
    int32_t cpr;      // 64-voxel chunks per x-row
    int32_t nchunks;  // cpr * gy * gz
};

__global__ __launch_bounds__(256) void weld_mask_kernel(const WeldParams w,
                                                        unsigned long long* __restrict__ masks,
                                                        int32_t* __restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * kWeldWaves + (threadIdx.x >> 6);
    if (c >= w.nchunks) return;
    const McParams& p = w.mc;
    const int row = c / w.cpr;  // y + gy*z
    const int x = (c % w.cpr) * 64 + lane;
    const int y = row % p.gy, z = row / p.gy;
    bool bx = false, by = false, bz = false;
    if (x < p.gx) {
        const int64_t i = (int64_t)x + (int64_t)p.gx * row;
        const bool n0 = p.f[i] < 0.0f;
        bx = x + 1 < p.gx && (p.f[i + 1] < 0.0f) != n0;
        by = y + 1 < p.gy && (p.f[i + p.gx] < 0.0f) != n0;
        bz = z + 1 < p.gz && (p.f[i + (int64_t)p.gx * p.gy] < 0.0f) != n0;
    }
    const unsigned long long mx = __ballot(bx), my = __ballot(by), mz = __ballot(bz);
    if (lane == 0) {
        masks[3 * (int64_t)c + 0] = mx;
        masks[3 * (int64_t)c + 1] = my;
        masks[3 * (int64_t)c + 2] = mz;
        counts[c] = __popcll(mx) + __popcll(my) + __popcll(mz);
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

__global__ __launch_bounds__(256) void weld_vertex_kernel(const WeldParams w,
                                                          const unsigned long long* __restrict__ masks,
                                                          const long long* __restrict__ base,
                                                          float* __restrict__ verts,
                                                          float* __restrict__ normals) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * kWeldWaves + (threadIdx.x >> 6);
    if (c >= w.nchunks) return;
    const McParams& p = w.mc;
    const unsigned long long m[3] = {masks[3 * (int64_t)c], masks[3 * (int64_t)c + 1],
                                     masks[3 * (int64_t)c + 2]};
    const unsigned long long below = (1ull << lane) - 1ull;
    long long idx = base[c] + __popcll(m[0] & below) + __popcll(m[1] & below) + __popcll(m[2] & below);
    const int row = c / w.cpr;
    const int x = (c % w.cpr) * 64 + lane;
    const int y = row % p.gy, z = row / p.gy;
    for (int a = 0; a < 3; ++a) {
        if (!((m[a] >> lane) & 1ull)) continue;
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
        ++idx;
    }
}

// Vertex index of local edge e of cube (x, y, z).
__device__ __forceinline__ int32_t weld_index(const WeldParams& w,
                                              const unsigned long long* __restrict__ masks,
                                              const long long* __restrict__ base, int x, int y,
                                              int z, int e) {
    // corner offsets packed as bits: dx = bit0, dy = bit1, dz = bit2
    // (0:000 1:001 2:101 3:100 4:010 5:011 6:111 7:110), as in edge_vertex
    constexpr unsigned k_off = (0u << 0) | (1u << 3) | (5u << 6) | (4u << 9) | (2u << 12) |
                               (3u << 15) | (7u << 18) | (6u << 21);
    const unsigned oa = (k_off >> (3 * c_edge[e][0])) & 7u, ob = (k_off >> (3 * c_edge[e][1])) & 7u;
    const int a = __builtin_ctz(oa ^ ob);  // the axis the edge runs along
    const unsigned lo = oa & ob;           // its lower end
    const int X = x + (int)(lo & 1u), Y = y + (int)((lo >> 1) & 1u), Z = z + (int)(lo >> 2);
    const int c = (Y + w.mc.gy * Z) * w.cpr + (X >> 6);
    const int l = X & 63;
    const unsigned long long m0 = masks[3 * (int64_t)c], m1 = masks[3 * (int64_t)c + 1],
                             m2 = masks[3 * (int64_t)c + 2];
    const unsigned long long below = (1ull << l) - 1ull;
    long long idx = base[c] + __popcll(m0 & below) + __popcll(m1 & below) + __popcll(m2 & below);
    if (a > 0) idx += (long long)((m0 >> l) & 1ull);
    if (a > 1) idx += (long long)((m1 >> l) & 1ull);
    return (int32_t)idx;
}

__global__ __launch_bounds__(256) void weld_row_count_kernel(const McParams p,
                                                             int32_t* __restrict__ row_count) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * kWeldWaves + (threadIdx.x >> 6);
    if (r >= p.rows) return;
    const int n = mc_row_triangles(p, r, lane);
    if (lane == 0) row_count[r] = n;
}

// Triangles of cube row r in the soup's order (one wavefront per row).
__global__ __launch_bounds__(256) void weld_emit_kernel(const WeldParams w,
                                                        const int32_t* __restrict__ row_count,
                                                        const long long* __restrict__ row_offset,
                                                        const unsigned long long* __restrict__ masks,
                                                        const long long* __restrict__ base,
                                                        int32_t* __restrict__ tris) {
    const int lane = threadIdx.x & 63;
    const McParams& p = w.mc;
    const int r = blockIdx.x * kWeldWaves + (threadIdx.x >> 6);
    if (r >= p.rows || row_count[r] == 0) return;
    const int y = r % p.cyr, z = r / p.cyr;
    long long t0 = row_offset[r];
    for (int x0 = 0; x0 < p.cx; x0 += 64) {
        const int x = x0 + lane;
        int c = 0, nt = 0;
        if (x < p.cx) {
            float v[8];
            load_corners(p, x, y, z, v);
            c = cube_class(v);
            nt = class_ntri(c);
        }
        const int incl = wave_inclusive_scan(nt, lane);
        long long t = t0 + (incl - nt);
        if (nt) {
            const uint64_t word = c_tri[c];
            for (int s = 0; s < nt; ++s, ++t)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int e = (int)((word >> (4 + 4 * (3 * s + k))) & 0xFull);
                    tris[3 * t + k] = weld_index(w, masks, base, x, y, z, e);
                }
        }
        t0 += __shfl(incl, 63, 64);
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
    const int nbc = (w.nchunks + kScanBlock - 1) / kScanBlock;
    const int nbr = (p.rows + kScanBlock - 1) / kScanBlock;
    const size_t o_masks = 0;
    const size_t o_counts = o_masks + align256(24 * (size_t)w.nchunks);
    const size_t o_base = o_counts + align256(4 * (size_t)w.nchunks);
    const size_t o_csums = o_base + align256(8 * (size_t)w.nchunks);
    const size_t o_rcount = o_csums + align256(8 * (size_t)(nbc + 1));
    const size_t o_roff = o_rcount + align256(4 * (size_t)p.rows);
    const size_t o_rsums = o_roff + align256(8 * (size_t)p.rows);
    const size_t need = o_rsums + align256(8 * (size_t)(nbr + 1));
    if (*scratch_bytes < need) {
        if (*scratch) (void)hipFree(*scratch);
        *scratch = nullptr;
        *scratch_bytes = 0;
        RECON_HIP_TRY(hipMalloc(scratch, need), "welded marching cubes scratch");
        *scratch_bytes = need;
    }
    char* sb = static_cast<char*>(*scratch);
    auto* masks = reinterpret_cast<unsigned long long*>(sb + o_masks);
    auto* counts = reinterpret_cast<int32_t*>(sb + o_counts);
    auto* base = reinterpret_cast<long long*>(sb + o_base);
    auto* csums = reinterpret_cast<long long*>(sb + o_csums);
    auto* rcount = reinterpret_cast<int32_t*>(sb + o_rcount);
    auto* roff = reinterpret_cast<long long*>(sb + o_roff);
    auto* rsums = reinterpret_cast<long long*>(sb + o_rsums);
    const int cblocks = (w.nchunks + kWeldWaves - 1) / kWeldWaves;
    const int rblocks = (p.rows + kWeldWaves - 1) / kWeldWaves;
    hipLaunchKernelGGL(weld_mask_kernel, dim3(cblocks), dim3(256), 0, s, w, masks, counts);
    RECON_HIP_TRY(hipGetLastError(), "weld_mask_kernel launch");
    int rc = launch_exclusive_scan(counts, w.nchunks, base, csums, s);
    if (rc) return rc;
    hipLaunchKernelGGL(weld_row_count_kernel, dim3(rblocks), dim3(256), 0, s, p, rcount);
    RECON_HIP_TRY(hipGetLastError(), "weld_row_count_kernel launch");
    rc = launch_exclusive_scan(rcount, p.rows, roff, rsums, s);
    if (rc) return rc;
    long long totals[2] = {0, 0};
    RECON_HIP_TRY(hipMemcpyAsync(&totals[0], csums + nbc, 8, hipMemcpyDeviceToHost, s), "vertex total");
    RECON_HIP_TRY(hipMemcpyAsync(&totals[1], rsums + nbr, 8, hipMemcpyDeviceToHost, s), "triangle total");
    RECON_HIP_TRY(hipStreamSynchronize(s), "welded marching cubes counts");
    *nv_out = totals[0];
    *nt_out = totals[1];
    if (!verts && !normals && !tris) return RECON_OK;
    if (totals[0] > INT32_MAX)
        return fail(RECON_ERR_RANGE, "%s: %lld vertices exceed int32 indices", fn, totals[0]);
    if ((verts || normals) && totals[0] > vcap)
        return fail(RECON_ERR_CAPACITY, "%s: %lld vertices do not fit capacity %lld", fn, totals[0],
                    (long long)vcap);
    if (tris && totals[1] > tcap)
        return fail(RECON_ERR_CAPACITY, "%s: %lld triangles do not fit capacity %lld", fn, totals[1],
                    (long long)tcap);
    if ((verts || normals) && totals[0] > 0) {
        hipLaunchKernelGGL(weld_vertex_kernel, dim3(cblocks), dim3(256), 0, s, w, masks, base, verts,
                           normals);
        RECON_HIP_TRY(hipGetLastError(), "weld_vertex_kernel launch");
    }
    if (tris && totals[1] > 0) {
        hipLaunchKernelGGL(weld_emit_kernel, dim3(rblocks), dim3(256), 0, s, w, rcount, roff, masks,
                           base, tris);
        RECON_HIP_TRY(hipGetLastError(), "weld_emit_kernel launch");
    }
    RECON_HIP_TRY(hipStreamSynchronize(s), "welded marching cubes");
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
