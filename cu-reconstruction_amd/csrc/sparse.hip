// sparse.hip -- sparse TSDF volume for gfx950: 8x8x8-voxel bricks found
// through a hash table, allocated along each frame's truncation band.
//
// No counterpart on the reference's path: its TSDFVolume is a dense grid
// (tsdf_volume.cu:194-254) and its only sparse structure is an unfinished
// octree prototype (octree.cu).  Here the lattice is the dense volume's --
// voxel v (any int32 triple, |v| < 2^23) has centre ((v + 0.5f) * vs) + off,
// the same float32 operations as tsdf_volume.cu:117-119 -- and a voxel holds
// D and W in the brick b = floor(v / 8) that contains it.  Unallocated voxels
// read as the fill value (D = truncation, W = 0).
//
// One integrate call (recon_sparse_integrate):
//   sparse_mark_kernel      counter snapshot (the first brick this call adds)
//   sparse_alloc_kernel     one lane per pixel with a valid depth d: the
//                           world box of the pixel's frustum slab between
//                           d - trunc and d + trunc (its eight corners), one
//                           voxel wider on every side, and every brick it
//                           meets is inserted (open addressing, linear
//                           probing, bounded, atomicCAS on the key)
//   sparse_init_kernel      new bricks := (truncation, 0)
//   sparse_integrate_kernel one wavefront per allocated brick, one x-row of
//                           8 voxels per lane: the dense update rule of
//                           integrate_device.hpp, voxel for voxel
// Every voxel the dense integrate would update inside the band
// (-trunc <= sdf <= trunc) lies in the slab of the pixel it projects to, so
// its brick is allocated before it is integrated; with every brick already
// allocated the result equals the dense integrate bit for bit.  Voxels of
// unallocated bricks miss only free-space (sdf > trunc) observations.
#pragma clang fp contract(off)

#include "recon_common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace recon {

constexpr int kBrick = 8;                        // voxels per brick edge
constexpr int kBrickVoxels = kBrick * kBrick * kBrick;
constexpr unsigned long long kEmptySlot = ~0ull;
constexpr int kMaxBricksPerAxis = 8;             // per pixel slab; more sets the overflow flag
constexpr int kCoordBits = 21;                   // brick coordinate range [-2^20, 2^20)
// Linear probing stops after kMaxProbes slots, on insertion and lookup
// alike, so a key is always within kMaxProbes of its hash.  The table has
// twice the pool's slots and stops taking keys when the pool is full, so it
// stays at most about half full and the bound is almost never reached; when
// it is, the brick is dropped and flagged.
constexpr int kMaxProbes = 128;

struct SparseParams {
    float vs[3], off[3];
    float fx, fy, cx, cy;
    float trunc, maxw, maxd;
    int64_t capacity;                 // bricks in the pool
    uint64_t mask;                    // hash table size - 1 (power of two)
    unsigned long long* keys;         // table keys, kEmptySlot when free
    int32_t* vals;                    // table values: brick index
    int4* coords;                     // brick coordinates, per brick
    float* D;                         // capacity * 512
    float* W;
    unsigned* counter;                // bricks handed out (may exceed capacity)
    unsigned* first_new;              // counter at the start of the call
    unsigned* overflow;               // bit 0: pool full, bit 1: slab too wide, bit 2: coordinate
                                      // range, bit 3: probe limit
    int32_t blo[3], bhi[3];           // allocation bounds in bricks (inclusive)
};

__host__ __device__ inline unsigned long long brick_key(int bx, int by, int bz) {
    const unsigned long long m = (1ull << kCoordBits) - 1ull;
    return (((unsigned long long)bx & m) << (2 * kCoordBits)) |
           (((unsigned long long)by & m) << kCoordBits) | ((unsigned long long)bz & m);
}

__host__ __device__ inline uint64_t hash_key(unsigned long long k) {
    // murmur3 64-bit finaliser
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ int floor_div8(int v) { return v >= 0 ? v >> 3 : -((-v + 7) >> 3); }


// Brick index of (bx, by, bz), or -1.
__device__ __forceinline__ int32_t find_brick(const SparseParams& p, int bx, int by, int bz) {
    const unsigned long long key = brick_key(bx, by, bz);
    uint64_t h = hash_key(key) & p.mask;
    for (uint64_t n = 0; n < (uint64_t)kMaxProbes && n <= p.mask; ++n) {
        const unsigned long long k = p.keys[h];
        if (k == key) return p.vals[h];
        if (k == kEmptySlot) return -1;
        h = (h + 1) & p.mask;
    }
    return -1;
}

__device__ __forceinline__ void insert_brick(const SparseParams& p, int bx, int by, int bz) {
    const unsigned long long key = brick_key(bx, by, bz);
    uint64_t h = hash_key(key) & p.mask;
    for (uint64_t n = 0; n < (uint64_t)kMaxProbes && n <= p.mask; ++n) {
        unsigned long long k = p.keys[h];
        if (k == key) return;
        if (k == kEmptySlot) {
            // a full pool takes no more keys (a racing few may still pass)
            if ((int64_t)*p.counter >= p.capacity) {
                atomicOr(p.overflow, 1u);
                return;
            }
            k = atomicCAS(p.keys + h, kEmptySlot, key);
            if (k == kEmptySlot) {
                const unsigned idx = atomicAdd(p.counter, 1u);
                if ((int64_t)idx < p.capacity) {
                    p.vals[h] = (int32_t)idx;
                    p.coords[idx] = make_int4(bx, by, bz, 0);
                } else {
                    p.vals[h] = -1;
                    atomicOr(p.overflow, 1u);
                }
                return;
            }
            if (k == key) return;
        }
        h = (h + 1) & p.mask;
    }
    atomicOr(p.overflow, 8u);
}

__global__ void sparse_mark_kernel(const SparseParams p) {
    if (threadIdx.x == 0) *p.first_new = *p.counter;
}

__global__ __launch_bounds__(256) void sparse_alloc_kernel(const SparseParams p, const FrameParams F,
                                                           int32_t cols, int32_t rows) {
    const int u = blockIdx.x * 64 + (threadIdx.x & 63);
    const int v = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (u >= cols || v >= rows) return;
    const float d = F.depth[(int64_t)v * F.pitch + u];
    if (!(d > 0.0f && d <= p.maxd)) return;
    const float z0 = d - p.trunc, z1 = d + p.trunc;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float su = (float)(u + (c & 1)), sv = (float)(v + ((c >> 1) & 1));
        const float z = (c & 4) ? z1 : z0;
        // camera point of pixel corner (su, sv) at depth z, then world = R^T (pc - t)
        const float xc = ((su - p.cx) * z) / p.fx;
        const float yc = ((sv - p.cy) * z) / p.fy;
        const float qx = xc - F.t[0], qy = yc - F.t[1], qz = z - F.t[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float w = ((F.R[k] * qx + F.R[3 + k] * qy) + F.R[6 + k] * qz);
            lo[k] = fminf(lo[k], w);
            hi[k] = fmaxf(hi[k], w);
        }
    }
    int b0[3], b1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        // voxel index range, one voxel wider on each side (covers rounding)
        const float a = floorf((lo[k] - p.off[k]) / p.vs[k]) - 1.0f;
        const float b = floorf((hi[k] - p.off[k]) / p.vs[k]) + 1.0f;
        if (!(a > -8388608.0f && b < 8388608.0f)) {  // |v| < 2^23
            atomicOr(p.overflow, 4u);
            return;
        }
        b0[k] = floor_div8((int)a);
        b1[k] = floor_div8((int)b);
        if (b1[k] - b0[k] >= kMaxBricksPerAxis) {
            atomicOr(p.overflow, 2u);
            b1[k] = b0[k] + kMaxBricksPerAxis - 1;
        }
        // allocation bounds (recon_sparse_set_bounds)
        b0[k] = max(b0[k], p.blo[k]);
        b1[k] = min(b1[k], p.bhi[k]);
    }
    for (int bz = b0[2]; bz <= b1[2]; ++bz)
        for (int by = b0[1]; by <= b1[1]; ++by)
            for (int bx = b0[0]; bx <= b1[0]; ++bx) insert_brick(p, bx, by, bz);
}

__global__ __launch_bounds__(256) void sparse_init_kernel(const SparseParams p) {
    const unsigned first = *p.first_new;
    const unsigned end = (unsigned)min((int64_t)*p.counter, p.capacity);
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)first * kBrickVoxels + blockIdx.x * 256 + threadIdx.x;
         i < (int64_t)end * kBrickVoxels; i += stride) {
        p.D[i] = p.trunc;
        p.W[i] = 0.0f;
    }
}

// Conservative visibility of a brick for one frame (wave-uniform result).
// Lanes 0..7 take the brick's corner voxel centres to the camera; a voxel
// can only pass the per-voxel tests when its camera point p satisfies
//   z > 0,  fx*x + (cx+1)*z > 0,  (cols-cx)*z - fx*x > 0,
//   fy*y + (cy+1)*z > 0,  (rows-cy)*z - fy*y > 0,  z <= maxd + trunc,
// all linear in p, so when every corner violates one of them by more than a
// relative 1e-3 every voxel centre of the brick (inside the corners' hull)
// does too, and the brick is skipped.
__device__ __forceinline__ bool brick_may_update(const SparseParams& p, const FrameParams& F,
                                                 const int4 bc, float fcols, float frows, int lane) {
    unsigned out = 0;  // bit k: this corner violates constraint k
    if (lane < 8) {
        const int v[3] = {bc.x * kBrick + ((lane & 1) ? kBrick - 1 : 0),
                          bc.y * kBrick + ((lane & 2) ? kBrick - 1 : 0),
                          bc.z * kBrick + ((lane & 4) ? kBrick - 1 : 0)};
        float c[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) c[k] = voxel_centre(v[k], p.vs[k], p.off[k]);
        float q[3];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            q[k] = ((F.R[3 * k] * c[0] + F.R[3 * k + 1] * c[1]) + F.R[3 * k + 2] * c[2]) + F.t[k];
        const float x = q[0], y = q[1], z = q[2];
        const float ax = fabsf(p.fx * x), ay = fabsf(p.fy * y), az = fabsf(z);
        const float e = 1e-3f;
        if (z < -e * (az + 1e-6f)) out |= 1u;
        if (p.fx * x + (p.cx + 1.0f) * z < -e * (ax + fabsf(p.cx + 1.0f) * az + 1e-6f)) out |= 2u;
        if ((fcols - p.cx) * z - p.fx * x < -e * (ax + fabsf(fcols - p.cx) * az + 1e-6f)) out |= 4u;
        if (p.fy * y + (p.cy + 1.0f) * z < -e * (ay + fabsf(p.cy + 1.0f) * az + 1e-6f)) out |= 8u;
        if ((frows - p.cy) * z - p.fy * y < -e * (ay + fabsf(frows - p.cy) * az + 1e-6f)) out |= 16u;
        if (z > (p.maxd + p.trunc) * (1.0f + e)) out |= 32u;
    } else {
        out = 63u;
    }
    // a constraint every corner violates
    unsigned all = out;
#pragma unroll
    for (int d = 1; d < 8; d <<= 1) all &= __shfl_xor(all, d, 64);
    return __builtin_amdgcn_readfirstlane(all) == 0u;
}

// One wavefront per brick (grid-stride over the allocated bricks); lane l
// holds the x-row (y, z) = (l % 8, l / 8) of the brick.
__global__ __launch_bounds__(256) void sparse_integrate_kernel(const SparseParams p,
                                                               const FrameParams F, float fcols,
                                                               float frows) {
    const int lane = threadIdx.x & 63;
    const unsigned n = (unsigned)min((int64_t)__builtin_amdgcn_readfirstlane(*p.counter), p.capacity);
    const unsigned waves = gridDim.x * 4;
    for (unsigned b = blockIdx.x * 4 + (threadIdx.x >> 6); b < n; b += waves) {
        const int4 bc = p.coords[b];
        if (!brick_may_update(p, F, bc, fcols, frows, lane)) continue;
        const int y = bc.y * kBrick + (lane & 7), z = bc.z * kBrick + (lane >> 3);
        const float cyw = voxel_centre(y, p.vs[1], p.off[1]);
        const float czw = voxel_centre(z, p.vs[2], p.off[2]);
        const float rxy = F.R[1] * cyw, ryy = F.R[4] * cyw, rzy = F.R[7] * cyw;
        bool pass[kBrick];
        float tsdf[kBrick];
        bool any = false;
#pragma unroll
        for (int i = 0; i < kBrick; ++i) {
            const float cxw = voxel_centre(bc.x * kBrick + i, p.vs[0], p.off[0]);
            const float px = ((F.R[0] * cxw + rxy) + F.R[2] * czw) + F.t[0];
            const float py = ((F.R[3] * cxw + ryy) + F.R[5] * czw) + F.t[1];
            const float pz = ((F.R[6] * cxw + rzy) + F.R[8] * czw) + F.t[2];
            int u = 0, v = 0;
            const bool ok = pz > 0.0f && pixel_index(px, p.fx, pz, p.cx, fcols, u) &&
                            pixel_index(py, p.fy, pz, p.cy, frows, v);
            const float dep = F.depth[ok ? (int64_t)v * F.pitch + u : 0];
            const float sdf = dep - pz;
            pass[i] = ok && !(dep <= 0.0f || dep > p.maxd) && (sdf >= -p.trunc);
            tsdf[i] = sdf > 0.0f ? fminf(sdf, p.trunc) : sdf;
            any = any || pass[i];
        }
        if (!any) continue;
        float4* Dr = reinterpret_cast<float4*>(p.D + (int64_t)b * kBrickVoxels + lane * kBrick);
        float4* Wr = reinterpret_cast<float4*>(p.W + (int64_t)b * kBrickVoxels + lane * kBrick);
        float4 d4[2] = {Dr[0], Dr[1]}, w4[2] = {Wr[0], Wr[1]};
        float* d = reinterpret_cast<float*>(d4);
        float* w = reinterpret_cast<float*>(w4);
#pragma unroll
        for (int i = 0; i < kBrick; ++i) {
            if (!pass[i]) continue;
            // tsdf_volume.cu:178-186 with the weight cap, as the dense path
            const float nw = w[i] + 1.0f;
            d[i] = ((d[i] * w[i]) + tsdf[i]) / nw;
            w[i] = nw < p.maxw ? nw : p.maxw;
        }
        Dr[0] = d4[0];
        Dr[1] = d4[1];
        Wr[0] = w4[0];
        Wr[1] = w4[1];
    }
}

__global__ __launch_bounds__(256) void sparse_alloc_box_kernel(const SparseParams p, int4 b0, int4 nb) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)nb.x * nb.y * nb.z;
    if (i >= total) return;
    const int bx = b0.x + (int)(i % nb.x), by = b0.y + (int)((i / nb.x) % nb.y),
              bz = b0.z + (int)(i / ((int64_t)nb.x * nb.y));
    insert_brick(p, bx, by, bz);
}

// Dense box [o, o + size) of the sparse volume; unallocated voxels give the
// fill value.
__global__ __launch_bounds__(256) void sparse_export_kernel(const SparseParams p, int4 o, int4 size,
                                                            float* __restrict__ D,
                                                            float* __restrict__ W) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)size.x * size.y * size.z;
    if (i >= total) return;
    const int x = o.x + (int)(i % size.x), y = o.y + (int)((i / size.x) % size.y),
              z = o.z + (int)(i / ((int64_t)size.x * size.y));
    const int32_t b = find_brick(p, floor_div8(x), floor_div8(y), floor_div8(z));
    float dv = p.trunc, wv = 0.0f;
    if (b >= 0) {
        const int64_t j = (int64_t)b * kBrickVoxels + ((z & 7) * 8 + (y & 7)) * 8 + (x & 7);
        dv = p.D[j];
        wv = p.W[j];
    }
    if (D) D[i] = dv;
    if (W) W[i] = wv;
}

__global__ __launch_bounds__(256) void sparse_clear_kernel(unsigned long long* __restrict__ keys,
                                                           uint64_t n, unsigned* counters) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) keys[i] = kEmptySlot;
    if (i < 3) counters[i] = 0u;
}

}  // namespace recon

using namespace recon;

struct recon_sparse {
    recon_tsdf_info info;
    recon_tsdf_config cfg;
    SparseParams p;
    void* mem = nullptr;  // one allocation: table, bricks, counters
    int num_cus = 256;
};

extern "C" int recon_sparse_create(const recon_tsdf_config* cfg, int64_t brick_capacity,
                                   recon_sparse** out) {
    clear_error();
    const char* fn = "recon_sparse_create";
    if (!out) return fail(RECON_ERR_NULL, "%s: out is NULL", fn);
    *out = nullptr;
    recon_tsdf_info info;
    int rc = describe(cfg, &info);
    if (rc) return rc;
    if (info.weight_type != RECON_WEIGHT_BY_DEPTH)
        return fail(RECON_ERR_CONFIG, "%s: sparse volumes integrate with WEIGHT_BY_DEPTH only", fn);
    if (brick_capacity < 1 || brick_capacity > (int64_t(1) << 24))
        return fail(RECON_ERR_RANGE, "%s: brick capacity %lld outside 1..2^24", fn,
                    (long long)brick_capacity);
    rc = require_gfx950();
    if (rc) return rc;
    uint64_t table = 1;
    while (table < 2 * (uint64_t)brick_capacity) table <<= 1;
    const size_t bytes_keys = 8 * table, bytes_vals = 4 * table;
    const size_t bytes_coords = 16 * (size_t)brick_capacity;
    const size_t bytes_vox = 4 * (size_t)kBrickVoxels * (size_t)brick_capacity;
    auto a256 = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t o_vals = a256(bytes_keys), o_coords = o_vals + a256(bytes_vals),
                 o_D = o_coords + a256(bytes_coords), o_W = o_D + a256(bytes_vox),
                 o_cnt = o_W + a256(bytes_vox), total = o_cnt + 256;
    recon_sparse* h = new recon_sparse();
    h->cfg = *cfg;
    h->info = info;
    if (hipMalloc(&h->mem, total) != hipSuccess) {
        delete h;
        return fail(RECON_ERR_DEVICE, "%s: cannot allocate %zu bytes", fn, total);
    }
    char* m = static_cast<char*>(h->mem);
    SparseParams& p = h->p;
    for (int k = 0; k < 3; ++k) {
        p.vs[k] = info.voxel_size[k];
        p.off[k] = info.global_offset[k];
    }
    p.fx = cfg->focal_x;
    p.fy = cfg->focal_y;
    p.cx = cfg->cx;
    p.cy = cfg->cy;
    p.trunc = info.truncation;
    p.maxw = cfg->max_weight;
    p.maxd = info.max_depth;
    p.capacity = brick_capacity;
    p.mask = table - 1;
    p.keys = reinterpret_cast<unsigned long long*>(m);
    p.vals = reinterpret_cast<int32_t*>(m + o_vals);
    p.coords = reinterpret_cast<int4*>(m + o_coords);
    p.D = reinterpret_cast<float*>(m + o_D);
    p.W = reinterpret_cast<float*>(m + o_W);
    for (int k = 0; k < 3; ++k) {
        p.blo[k] = -(1 << (kCoordBits - 1));
        p.bhi[k] = (1 << (kCoordBits - 1)) - 1;
    }
    p.counter = reinterpret_cast<unsigned*>(m + o_cnt);
    p.first_new = p.counter + 1;
    p.overflow = p.counter + 2;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) h->num_cus = prop.multiProcessorCount;
    }
    hipLaunchKernelGGL(sparse_clear_kernel, dim3((unsigned)((table + 255) / 256)), dim3(256), 0,
                       nullptr, p.keys, table, p.counter);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(h->mem);
        delete h;
        return fail(RECON_ERR_DEVICE, "%s: clear failed", fn);
    }
    *out = h;
    return RECON_OK;
}

extern "C" void recon_sparse_destroy(recon_sparse* h) {
    if (!h) return;
    if (h->mem) (void)hipFree(h->mem);
    delete h;
}

extern "C" int recon_sparse_reset(recon_sparse* h, recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_sparse_reset: handle is NULL");
    const uint64_t table = h->p.mask + 1;
    hipLaunchKernelGGL(sparse_clear_kernel, dim3((unsigned)((table + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, h->p.keys, table, h->p.counter);
    RECON_HIP_TRY(hipGetLastError(), "sparse_clear_kernel launch");
    return RECON_OK;
}

static int sparse_new_bricks(recon_sparse* h, hipStream_t s) {
    hipLaunchKernelGGL(sparse_init_kernel, dim3((unsigned)(h->num_cus * 8)), dim3(256), 0, s, h->p);
    RECON_HIP_TRY(hipGetLastError(), "sparse_init_kernel launch");
    return RECON_OK;
}

extern "C" int recon_sparse_integrate(recon_sparse* h, const recon_depth_view* depth,
                                      const double camera_pose[7], recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_sparse_integrate";
    if (!h) return fail(RECON_ERR_NULL, "%s: handle is NULL", fn);
    if (!depth || !depth->data) return fail(RECON_ERR_NULL, "%s: depth is NULL", fn);
    if (depth->rows <= 0 || depth->cols <= 0)
        return fail(RECON_ERR_SIZE, "%s: depth size %dx%d must be positive", fn, depth->cols,
                    depth->rows);
    if (depth->step < 4 * (size_t)depth->cols || depth->step % 4 != 0)
        return fail(RECON_ERR_SIZE, "%s: depth step %zu invalid", fn, depth->step);
    if ((depth->step / 4) * (size_t)depth->rows >= (size_t(1) << 31) || depth->rows >= (1 << 18))
        return fail(RECON_ERR_RANGE, "%s: depth image too large", fn);
    if (!camera_pose) return fail(RECON_ERR_NULL, "%s: pose is NULL", fn);
    if (!pose_is_finite(camera_pose)) return fail(RECON_ERR_CONFIG, "%s: non-finite pose", fn);
    hipStream_t s = (hipStream_t)stream;
    FrameParams F;
    pose_to_matrix_f32(camera_pose, F.R, F.t);
    F.depth = depth->data;
    F.pitch = (int32_t)(depth->step / 4);
    hipLaunchKernelGGL(sparse_mark_kernel, dim3(1), dim3(64), 0, s, h->p);
    RECON_HIP_TRY(hipGetLastError(), "sparse_mark_kernel launch");
    hipLaunchKernelGGL(sparse_alloc_kernel, dim3((depth->cols + 63) / 64, (depth->rows + 3) / 4),
                       dim3(256), 0, s, h->p, F, depth->cols, depth->rows);
    RECON_HIP_TRY(hipGetLastError(), "sparse_alloc_kernel launch");
    int rc = sparse_new_bricks(h, s);
    if (rc) return rc;
    hipLaunchKernelGGL(sparse_integrate_kernel, dim3((unsigned)(h->num_cus * 8)), dim3(256), 0, s,
                       h->p, F, (float)depth->cols, (float)depth->rows);
    RECON_HIP_TRY(hipGetLastError(), "sparse_integrate_kernel launch");
    return RECON_OK;
}

extern "C" int recon_sparse_allocate_box(recon_sparse* h, const int32_t lo[3], const int32_t hi[3],
                                         recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_sparse_allocate_box";
    if (!h || !lo || !hi) return fail(RECON_ERR_NULL, "%s: null argument", fn);
    int b0[3], nb[3];
    int64_t total = 1;
    for (int k = 0; k < 3; ++k) {
        if (hi[k] <= lo[k]) return fail(RECON_ERR_SIZE, "%s: empty box", fn);
        if (lo[k] < -(1 << 23) || hi[k] > (1 << 23)) return fail(RECON_ERR_RANGE, "%s: box beyond 2^23", fn);
        const int a = lo[k] >= 0 ? lo[k] / 8 : -((-lo[k] + 7) / 8);
        const int hm = hi[k] - 1;
        const int b = hm >= 0 ? hm / 8 : -((-hm + 7) / 8);
        b0[k] = a;
        nb[k] = b - a + 1;
        total *= nb[k];
    }
    if (total > h->p.capacity)
        return fail(RECON_ERR_CAPACITY, "%s: %lld bricks exceed the capacity %lld", fn,
                    (long long)total, (long long)h->p.capacity);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(sparse_mark_kernel, dim3(1), dim3(64), 0, s, h->p);
    RECON_HIP_TRY(hipGetLastError(), "sparse_mark_kernel launch");
    hipLaunchKernelGGL(sparse_alloc_box_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       h->p, make_int4(b0[0], b0[1], b0[2], 0), make_int4(nb[0], nb[1], nb[2], 0));
    RECON_HIP_TRY(hipGetLastError(), "sparse_alloc_box_kernel launch");
    return sparse_new_bricks(h, s);
}

extern "C" int recon_sparse_set_bounds(recon_sparse* h, const int32_t lo[3], const int32_t hi[3]) {
    clear_error();
    const char* fn = "recon_sparse_set_bounds";
    if (!h) return fail(RECON_ERR_NULL, "%s: handle is NULL", fn);
    for (int k = 0; k < 3; ++k) {
        if (!lo || !hi) {
            h->p.blo[k] = -(1 << (kCoordBits - 1));
            h->p.bhi[k] = (1 << (kCoordBits - 1)) - 1;
            continue;
        }
        if (hi[k] <= lo[k]) return fail(RECON_ERR_SIZE, "%s: empty box", fn);
        if (lo[k] < -(1 << 23) || hi[k] > (1 << 23)) return fail(RECON_ERR_RANGE, "%s: box beyond 2^23", fn);
    }
    if (lo && hi)
        for (int k = 0; k < 3; ++k) {
            h->p.blo[k] = lo[k] >= 0 ? lo[k] / 8 : -((-lo[k] + 7) / 8);
            const int hm = hi[k] - 1;
            h->p.bhi[k] = hm >= 0 ? hm / 8 : -((-hm + 7) / 8);
        }
    return RECON_OK;
}

extern "C" int recon_sparse_stats(recon_sparse* h, int64_t* num_bricks, uint32_t* overflow,
                                  recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_sparse_stats: handle is NULL");
    unsigned c[3] = {0, 0, 0};
    hipStream_t s = (hipStream_t)stream;
    RECON_HIP_TRY(hipMemcpyAsync(c, h->p.counter, sizeof c, hipMemcpyDeviceToHost, s), "sparse stats");
    RECON_HIP_TRY(hipStreamSynchronize(s), "sparse stats synchronize");
    if (num_bricks) *num_bricks = std::min<int64_t>(c[0], h->p.capacity);
    if (overflow) *overflow = c[2];
    return RECON_OK;
}

extern "C" int recon_sparse_export(recon_sparse* h, const int32_t origin[3], const int32_t size[3],
                                   float* D, float* W, recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_sparse_export";
    if (!h || !origin || !size) return fail(RECON_ERR_NULL, "%s: null argument", fn);
    if (!D && !W) return fail(RECON_ERR_NULL, "%s: no output", fn);
    int64_t total = 1;
    for (int k = 0; k < 3; ++k) {
        if (size[k] < 1) return fail(RECON_ERR_SIZE, "%s: size[%d] < 1", fn, k);
        if ((int64_t)origin[k] < -(1 << 23) || (int64_t)origin[k] + size[k] > (1 << 23))
            return fail(RECON_ERR_RANGE, "%s: box beyond 2^23", fn);
        total *= size[k];
    }
    if (total >= (int64_t(1) << 31)) return fail(RECON_ERR_RANGE, "%s: box too large", fn);
    hipLaunchKernelGGL(sparse_export_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, h->p, make_int4(origin[0], origin[1], origin[2], 0),
                       make_int4(size[0], size[1], size[2], 0), D, W);
    RECON_HIP_TRY(hipGetLastError(), "sparse_export_kernel launch");
    return RECON_OK;
}

extern "C" int recon_sparse_download_bricks(recon_sparse* h, int32_t* coords, float* D, float* W,
                                            int64_t capacity, int64_t* num_bricks,
                                            recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_sparse_download_bricks";
    if (!h || !num_bricks) return fail(RECON_ERR_NULL, "%s: null argument", fn);
    int64_t n = 0;
    int rc = recon_sparse_stats(h, &n, nullptr, stream);
    if (rc) return rc;
    *num_bricks = n;
    if (!coords && !D && !W) return RECON_OK;
    if (n > capacity) return fail(RECON_ERR_CAPACITY, "%s: %lld bricks do not fit", fn, (long long)n);
    hipStream_t s = (hipStream_t)stream;
    if (coords && n) {
        std::vector<int4> c((size_t)n);
        RECON_HIP_TRY(hipMemcpyAsync(c.data(), h->p.coords, 16 * (size_t)n, hipMemcpyDeviceToHost, s),
                      "brick coordinates");
        RECON_HIP_TRY(hipStreamSynchronize(s), "brick coordinates synchronize");
        for (int64_t i = 0; i < n; ++i) {
            coords[3 * i] = c[i].x;
            coords[3 * i + 1] = c[i].y;
            coords[3 * i + 2] = c[i].z;
        }
    }
    const size_t bytes = 4 * (size_t)kBrickVoxels * (size_t)n;
    if (D && n) RECON_HIP_TRY(hipMemcpyAsync(D, h->p.D, bytes, hipMemcpyDeviceToHost, s), "brick D");
    if (W && n) RECON_HIP_TRY(hipMemcpyAsync(W, h->p.W, bytes, hipMemcpyDeviceToHost, s), "brick W");
    RECON_HIP_TRY(hipStreamSynchronize(s), "brick download synchronize");
    return RECON_OK;
}
