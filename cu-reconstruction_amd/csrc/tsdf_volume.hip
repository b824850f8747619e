// tsdf_volume.hip -- TSDF volume lifecycle and the projective integrate
// kernels for gfx950.
//
// Replaces TSDFVolume (cu-Reconstruction modules/src/cuda/tsdf_volume.cu:
// 144-290).  The reference runs one thread per (y,z) voxel row looping over x
// (tsdf_volume.cu:151-163: neighbouring lanes are a row apart, so no access
// coalesces) and re-reads a 28-byte voxel-centre table per voxel.  Here one
// integrate call is three launches on the caller's stream:
//   1. depth_bounds_kernel  per 8x8 and 32x32 pixel block, the largest valid
//                           depth of each frame;
//   2. cull_kernel          one thread per wave-tile (32 x 8 x 2 voxels):
//                           the tile survives a frame if its voxel-centre box
//                           meets the frame's frustum and the depth bound of
//                           its image footprint is not closer than the box
//                           minus the truncation; each workgroup compacts its
//                           survivors (with their frame masks) into its own
//                           segment of the list and writes their count;
//   3. integrate_list_kernel  a grid sized to the resident wave slots walks
//                           the list, one wave-tile per wavefront: lane l owns
//                           voxels x..x+3 of row y for two z, so a wave moves
//                           aligned 128-byte row segments of D and W with
//                           dwordx4 accesses, and only lanes with a voxel that
//                           passes read (and write) D and W.
// Voxel centres are computed from the index (no table traffic).  Up to
// kMaxFusedFrames frames can go through one sweep: each voxel sees the frames
// in order, so the result equals one sweep per frame.  The culling is
// conservative (margins far above float32 rounding), so every voxel the
// reference updates is updated here with the same float32 operations.
#pragma clang fp contract(off)

#include "integrate_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace recon {

using Shape = DefaultShape;

__global__ __launch_bounds__(256) void depth_bounds_kernel(const BoundsJob job, size_t frame_stride,
                                                           size_t bounds_stride) {
    __shared__ int s0[16];
    __shared__ int s1[4];
    BoundsJob j = job;
    j.depth += frame_stride * blockIdx.z;
    j.bound0 += bounds_stride * blockIdx.z;
    j.bound1 += bounds_stride * blockIdx.z;
    bounds_block(j, blockIdx.x, blockIdx.y, s0, s1);
}

// Footprints wider than this many 8x8 depth blocks skip the fine bound test
// (the tile is kept after the 32x32 test).
constexpr int kCullMaxFine = 24;

// Workgroup b tests tiles [b*seg_len, (b+1)*seg_len) (seg_len a multiple of
// 256, one tile per thread per pass), compacts the survivors into the same
// slots of the list in tile order and writes their number to counts[b].
template <int MAXF>
__global__ __launch_bounds__(256) void cull_kernel(const VolumeParams vp, const int nframes,
                                                   const FrameBlock<MAXF> fb,
                                                   ListEntry* __restrict__ list,
                                                   unsigned* __restrict__ counts,
                                                   const unsigned seg_len) {
    __shared__ unsigned s_wave[kThreads / 64];
    const uint32_t ntiles = (uint32_t)vp.tiles_x * vp.tiles_y * vp.tiles_z;
    const uint32_t first = blockIdx.x * seg_len;
    const uint32_t end = min(first + seg_len, ntiles);
    unsigned n = 0;
    for (uint32_t k = first; k < end; k += kThreads) {
        const uint32_t id = k + threadIdx.x;
        const uint64_t live =
            id < end ? tile_frames<Shape, FrameBlock<MAXF>, kCullMaxFine>(vp, nframes, fb, id) : 0ull;
        n += append_local(live, id, 0u, list + first + n, s_wave);
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = n;
}

// The list's entries in dense order i = 0 .. total-1 (segment by segment):
// each wavefront takes i = wave, wave + waves, ... (the grid is sized to the
// resident wave slots, so every wave-tile costs about the same and the device
// stays full until the list is done); lane j of a wave locates the list slot
// of the wave's j-th next entry, so one search serves 64 entries.
template <int MAXF, bool VEC, bool COUNT, bool COLOR = false, bool VARW = false>
__global__ __launch_bounds__(kThreads) void integrate_list_kernel(
    const VolumeParams vp, const int nframes, const FrameBlock<MAXF> fb,
    const ListEntry* __restrict__ list, const unsigned* __restrict__ counts, const int nseg,
    const unsigned seg_len, unsigned long long* counter, unsigned long long* touch_counter) {
    __shared__ unsigned s_pre[kMaxSegments + 1];
    __shared__ unsigned s_wave[kThreads / 64];
    __shared__ float rtab[kRcpTable];
    rtab[threadIdx.x] = 1.0f / (float)(threadIdx.x + 1);  // ordered by segment_offsets' barrier
    const unsigned n = __builtin_amdgcn_readfirstlane(segment_offsets(counts, nseg, s_pre, s_wave));
    const unsigned waves = gridDim.x * (kThreads / 64);
    const unsigned lane = threadIdx.x & 63;
    unsigned cnt = 0, touched = 0;
    for (unsigned i0 = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); i0 < n; i0 += 64 * waves) {
        const unsigned mine = i0 + lane * waves;
        const unsigned slot = mine < n ? segment_slot(s_pre, nseg, seg_len, mine) : 0u;
        const unsigned m = min(64u, (n - i0 + waves - 1) / waves);
        for (unsigned j = 0; j < m; ++j) {
            const unsigned e = __builtin_amdgcn_readfirstlane(__shfl(slot, (int)j, 64));
            const uint32_t tile = __builtin_amdgcn_readfirstlane(list[e].tile);
            const uint64_t frames =
                ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(list[e].frames >> 32)) << 32) |
                __builtin_amdgcn_readfirstlane((uint32_t)list[e].frames);
            int x0, y0, z0;
            tile_origin<Shape>(vp, tile, x0, y0, z0);
            integrate_tile<Shape, FrameBlock<MAXF>, VEC, COUNT, COLOR, VARW>(vp, nframes, fb, frames,
                                                                             x0, y0, z0, cnt,
                                                                             touched, rtab);
        }
    }
    if (COUNT && counter && cnt) atomicAdd(counter, (unsigned long long)cnt);
    if (COUNT && touch_counter && touched) atomicAdd(touch_counter, (unsigned long long)touched);
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ D, float* __restrict__ W,
                                                   long long n, float value_d) {
    fill4(D, W, 4ll * (blockIdx.x * (long long)blockDim.x + threadIdx.x), n, value_d);
}

// init_voxel_transofrmation (tsdf_volume.cu:106-127): packed 28-byte records
// {float3 translation = voxel centre; float4 quat_rotation = (0, 0, 0, 1)}.
__global__ __launch_bounds__(256) void transformation_voxels_kernel(const VolumeParams vp,
                                                                    float* __restrict__ out,
                                                                    long long n) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % vp.gx), y = (int)((i / vp.gx) % vp.gy), z = (int)(i / ((long long)vp.gx * vp.gy));
    float* r = out + 7 * i;
    r[0] = voxel_centre(x, vp.vs[0], vp.off[0]);
    r[1] = voxel_centre(y, vp.vs[1], vp.off[1]);
    r[2] = voxel_centre(z, vp.vs[2], vp.off[2]);
    r[3] = 0.0f;
    r[4] = 0.0f;
    r[5] = 0.0f;
    r[6] = 1.0f;
}

__global__ __launch_bounds__(256) void fill_color_kernel(float4* __restrict__ C, long long n) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) C[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

static int launch_fill(recon_tsdf* h, hipStream_t s) {
    const long long n = h->info.num_voxels;
    const long long blocks = (n + 4 * 256 - 1) / (4 * 256);
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, h->D, h->W, n,
                       h->info.truncation);
    RECON_HIP_TRY(hipGetLastError(), "fill_kernel launch");
    if (h->C) {
        hipLaunchKernelGGL(fill_color_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<float4*>(h->C), n);
        RECON_HIP_TRY(hipGetLastError(), "fill_color_kernel launch");
    }
    return RECON_OK;
}

static int check_depth(const recon_depth_view* depth, const char* fn) {
    if (!depth) return fail(RECON_ERR_NULL, "%s: depth view is NULL", fn);
    if (!depth->data) return fail(RECON_ERR_NULL, "%s: depth data is NULL", fn);
    if (depth->rows <= 0 || depth->cols <= 0)
        return fail(RECON_ERR_SIZE, "%s: depth size %dx%d must be positive", fn, depth->cols,
                    depth->rows);
    if (depth->step < 4 * (size_t)depth->cols || depth->step % 4 != 0)
        return fail(RECON_ERR_SIZE, "%s: depth step %zu must be a multiple of 4 and >= 4*cols", fn,
                    depth->step);
    if ((depth->step / 4) * (size_t)depth->rows >= (size_t(1) << 31))
        return fail(RECON_ERR_RANGE, "%s: depth image exceeds 2^31 floats", fn);
    return RECON_OK;
}

// Grows the handle's per-call scratch (depth bounds of kMaxFusedFrames frames,
// the cull list and its segment counts) to fit `vp`.  A buffer being replaced may
// still be read by queued kernels, so the stream is drained first.
static int ensure_scratch(recon_tsdf* h, const VolumeParams& vp, hipStream_t s) {
    const size_t per = (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
    const size_t need = per * kMaxFusedFrames;
    if (h->bounds_floats < need) {
        if (h->bounds) {
            RECON_HIP_TRY(hipStreamSynchronize(s), "scratch resize synchronize");
            (void)hipFree(h->bounds);
        }
        h->bounds = nullptr;
        h->bounds_floats = 0;
        RECON_HIP_TRY(hipMalloc((void**)&h->bounds, need * sizeof(float)), "depth bounds scratch");
        h->bounds_floats = need;
    }
    return ensure_cull_list(h, (size_t)vp.tiles_x * vp.tiles_y * vp.tiles_z, sizeof(ListEntry));
}

// Cull segments: 256 tiles each, or the multiple of 256 that keeps their
// number within kMaxSegments.
static void cull_segments(int ntiles, unsigned* seg_len, int* nseg) {
    const unsigned need = (unsigned)((ntiles + kMaxSegments - 1) / kMaxSegments);
    *seg_len = std::max<unsigned>(kThreads, (need + kThreads - 1) / kThreads * kThreads);
    *nseg = (int)((ntiles + *seg_len - 1) / *seg_len);
}

struct ListArgs {
    const ListEntry* list;
    const unsigned* counts;
    int nseg;
    unsigned seg_len;
};

// At most this many list-kernel workgroups (4 waves each) per CU: at 8 the
// single-frame kernel ran 72.5 us per 512^3 call against 69.9 at 5 (each
// wave's share of the list shrinks and more tiles compete for the caches).
constexpr int kListMaxBlocksPerCU = 5;

// Workgroups per CU of one instantiation of the list kernel: its resident
// slots (capped).  The grid never exceeds the resident slots, because every
// wave takes the same share of the list and a workgroup beyond them would
// start only when another has finished its share.
template <int MAXF, bool VEC, bool COUNT, bool COLOR, bool VARW>
static int list_occupancy() {
    static const int blocks = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &b, integrate_list_kernel<MAXF, VEC, COUNT, COLOR, VARW>, kThreads, 0) != hipSuccess ||
            b < 1)
            b = 1;
        return std::min(b, kListMaxBlocksPerCU);
    }();
    return blocks;
}

template <int MAXF, bool VEC, bool COUNT, bool COLOR>
static void launch_list(bool varw, int num_cus, hipStream_t s, const VolumeParams& vp, int nframes,
                        const FrameBlock<MAXF>& fb, const ListArgs& la, unsigned long long* ctr,
                        unsigned long long* tctr) {
    if (varw)
        hipLaunchKernelGGL((integrate_list_kernel<MAXF, VEC, COUNT, COLOR, true>),
                           dim3(num_cus * list_occupancy<MAXF, VEC, COUNT, COLOR, true>()),
                           dim3(kThreads), 0, s, vp, nframes, fb, la.list, la.counts, la.nseg,
                           la.seg_len, ctr, tctr);
    else
        hipLaunchKernelGGL((integrate_list_kernel<MAXF, VEC, COUNT, COLOR, false>),
                           dim3(num_cus * list_occupancy<MAXF, VEC, COUNT, COLOR, false>()),
                           dim3(kThreads), 0, s, vp, nframes, fb, la.list, la.counts, la.nseg,
                           la.seg_len, ctr, tctr);
}

// Depth bounds and cull of one integrate sweep over `nframes` frames (frame f
// at depth0 + f * stride_floats) on stream s, into `bounds` and the cull
// list `list` / `counts`; fills the frames' bound pointers and `la`.
template <int MAXF>
static int launch_prestage(const VolumeParams& vp, const float* depth0, size_t stride_floats,
                           int pitch, int nframes, FrameBlock<MAXF>& fb, float* bounds,
                           ListEntry* list, unsigned* counts, hipStream_t s, ListArgs* la) {
    const size_t per = (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
    for (int f = 0; f < nframes; ++f) {
        fb.f[f].bound0 = bounds + per * f;
        fb.f[f].bound1 = fb.f[f].bound0 + (size_t)vp.b0x * vp.b0y;
    }
    BoundsJob job;
    job.depth = depth0;
    job.bound0 = bounds;
    job.bound1 = bounds + (size_t)vp.b0x * vp.b0y;
    job.pitch = pitch;
    job.cols = vp.cols;
    job.rows = vp.rows;
    job.b0x = vp.b0x;
    job.b0y = vp.b0y;
    job.b1x = vp.b1x;
    job.b1y = vp.b1y;
    job.maxd = vp.maxd;
    job.vec = (reinterpret_cast<uintptr_t>(depth0) % 16 == 0 && pitch % 4 == 0 && stride_floats % 4 == 0)
                  ? 1 : 0;
    hipLaunchKernelGGL(depth_bounds_kernel, dim3(vp.b1x, vp.b1y, nframes), dim3(256), 0, s, job,
                       stride_floats, per);
    RECON_HIP_TRY(hipGetLastError(), "depth_bounds_kernel launch");
    const int ntiles = vp.tiles_x * vp.tiles_y * vp.tiles_z;
    la->list = list;
    la->counts = counts;
    cull_segments(ntiles, &la->seg_len, &la->nseg);
    hipLaunchKernelGGL((cull_kernel<MAXF>), dim3(la->nseg), dim3(256), 0, s, vp, nframes, fb, list,
                       counts, la->seg_len);
    RECON_HIP_TRY(hipGetLastError(), "cull_kernel launch");
    return RECON_OK;
}

// The list kernel of one sweep (after launch_prestage's cull) on stream s.
template <int MAXF>
static int launch_liststage(recon_tsdf* h, const VolumeParams& vp, int nframes,
                            const FrameBlock<MAXF>& fb, const ListArgs& la, hipStream_t s) {
    const bool vec = (vp.gx % 4) == 0;
    const bool varw = h->info.weight_type == RECON_WEIGHT_BY_VARIANCE;
    unsigned long long* ctr = h->update_counter;
    unsigned long long* tctr = h->touch_counter;
    bool color = false;
    if constexpr (MAXF == 1) color = fb.f[0].rgb && vp.C;  // colour integrate (no counting)
    if (color) {
        if constexpr (MAXF == 1) {
            if (vec)
                launch_list<1, true, false, true>(varw, h->num_cus, s, vp, nframes, fb, la, ctr, tctr);
            else
                launch_list<1, false, false, true>(varw, h->num_cus, s, vp, nframes, fb, la, ctr, tctr);
        }
    } else if (ctr || tctr) {
        if (vec)
            launch_list<MAXF, true, true, false>(varw, h->num_cus, s, vp, nframes, fb, la, ctr, tctr);
        else
            launch_list<MAXF, false, true, false>(varw, h->num_cus, s, vp, nframes, fb, la, ctr, tctr);
    } else {
        if (vec)
            launch_list<MAXF, true, false, false>(varw, h->num_cus, s, vp, nframes, fb, la, ctr, tctr);
        else
            launch_list<MAXF, false, false, false>(varw, h->num_cus, s, vp, nframes, fb, la, ctr, tctr);
    }
    RECON_HIP_TRY(hipGetLastError(), "integrate_list_kernel launch");
    return RECON_OK;
}

// The three launches of one integrate sweep over `nframes` frames (frame f at
// depth0 + f * stride_floats), all on stream s.
template <int MAXF>
static int launch_integrate(recon_tsdf* h, const VolumeParams& vp, const float* depth0,
                            size_t stride_floats, int pitch, int nframes, FrameBlock<MAXF>& fb,
                            hipStream_t s) {
    int rc = ensure_scratch(h, vp, s);
    if (rc) return rc;
    ListArgs la;
    rc = launch_prestage<MAXF>(vp, depth0, stride_floats, pitch, nframes, fb, h->bounds,
                               static_cast<ListEntry*>(h->list), h->list_count, s, &la);
    if (rc) return rc;
    return launch_liststage<MAXF>(h, vp, nframes, fb, la, s);
}

static VolumeParams volume_params(const recon_tsdf* h, int cols, int rows) {
    VolumeParams vp = h->vp;
    vp.cols = cols;
    vp.rows = rows;
    vp.fcols = (float)cols;
    vp.frows = (float)rows;
    vp.b0x = (cols + kBound0 - 1) / kBound0;
    vp.b0y = (rows + kBound0 - 1) / kBound0;
    vp.b1x = (cols + kBound1 - 1) / kBound1;
    vp.b1y = (rows + kBound1 - 1) / kBound1;
    return vp;
}

}  // namespace recon

using namespace recon;

extern "C" int recon_tsdf_create(const recon_tsdf_config* cfg, recon_tsdf** out) {
    clear_error();
    if (!out) return fail(RECON_ERR_NULL, "recon_tsdf_create: out is NULL");
    *out = nullptr;
    recon_tsdf_info info;
    int rc = describe(cfg, &info);
    if (rc) return rc;
    rc = require_gfx950();
    if (rc) return rc;
    int device = 0, num_cus = 0;
    RECON_HIP_TRY(hipGetDevice(&device), "hipGetDevice");
    RECON_HIP_TRY(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, device),
                  "hipDeviceGetAttribute");
    recon_tsdf* h = new recon_tsdf();
    h->device = device;
    h->num_cus = num_cus;
    h->cfg = *cfg;
    h->info = info;
    VolumeParams& vp = h->vp;
    std::memset(&vp, 0, sizeof(vp));
    vp.gx = info.grid[0];
    vp.gy = info.grid[1];
    vp.gz = info.grid[2];
    vp.tiles_x = (vp.gx + Shape::X - 1) / Shape::X;
    vp.tiles_y = (vp.gy + Shape::Y - 1) / Shape::Y;
    vp.tiles_z = (vp.gz + Shape::Z - 1) / Shape::Z;
    for (int i = 0; i < 3; ++i) {
        vp.vs[i] = info.voxel_size[i];
        vp.off[i] = info.global_offset[i];
    }
    vp.fx = info.intrinsics[0];
    vp.fy = info.intrinsics[1];
    vp.cx = info.intrinsics[2];
    vp.cy = info.intrinsics[3];
    vp.trunc = info.truncation;
    vp.maxw = info.max_weight;
    vp.maxd = info.max_depth;
    if (info.weight_type == RECON_WEIGHT_BY_VARIANCE) {
        for (int i = 0; i < 3; ++i) vp.noise[i] = info.noise_model[i];
        const float e1 = 1.0f - vp.noise[2];
        const float s1 = vp.noise[0] + (vp.noise[1] * (e1 * e1));
        vp.wref = s1 * s1;
    }
    const size_t bytes = sizeof(float) * (size_t)info.num_voxels;
    hipError_t e = hipMalloc((void**)&h->D, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&h->W, bytes);
    vp.D = h->D;
    vp.W = h->W;
    if (e != hipSuccess) {
        rc = hip_fail(e, "recon_tsdf_create: hipMalloc");
        recon_tsdf_destroy(h);
        return rc;
    }
    rc = launch_fill(h, nullptr);
    if (rc == RECON_OK) {
        e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = hip_fail(e, "recon_tsdf_create: synchronize");
    }
    if (rc) {
        recon_tsdf_destroy(h);
        return rc;
    }
    *out = h;
    return RECON_OK;
}

extern "C" void recon_tsdf_destroy(recon_tsdf* h) {
    if (!h) return;
    if (h->D) (void)hipFree(h->D);
    if (h->W) (void)hipFree(h->W);
    if (h->mc_scratch) (void)hipFree(h->mc_scratch);
    if (h->bounds) (void)hipFree(h->bounds);
    if (h->list) (void)hipFree(h->list);
    if (h->diff_scratch) (void)hipFree(h->diff_scratch);
    if (h->owner) (void)hipFree(h->owner);
    if (h->C) (void)hipFree(h->C);
    delete h;
}

extern "C" int recon_tsdf_reset(recon_tsdf* h, recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_reset: handle is NULL");
    return launch_fill(h, (hipStream_t)stream);
}

extern "C" int recon_tsdf_get_info(const recon_tsdf* h, recon_tsdf_info* info) {
    clear_error();
    if (!h || !info) return fail(RECON_ERR_NULL, "recon_tsdf_get_info: null argument");
    *info = h->info;
    return RECON_OK;
}

extern "C" float* recon_tsdf_distances(recon_tsdf* h) { return h ? h->D : nullptr; }
extern "C" float* recon_tsdf_weights(recon_tsdf* h) { return h ? h->W : nullptr; }

extern "C" int recon_tsdf_download(recon_tsdf* h, float* host_d, float* host_w,
                                   recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_download: handle is NULL");
    const size_t bytes = sizeof(float) * (size_t)h->info.num_voxels;
    hipStream_t s = (hipStream_t)stream;
    if (host_d) RECON_HIP_TRY(hipMemcpyAsync(host_d, h->D, bytes, hipMemcpyDeviceToHost, s), "download D");
    if (host_w) RECON_HIP_TRY(hipMemcpyAsync(host_w, h->W, bytes, hipMemcpyDeviceToHost, s), "download W");
    RECON_HIP_TRY(hipStreamSynchronize(s), "recon_tsdf_download: synchronize");
    return RECON_OK;
}

extern "C" int recon_tsdf_upload(recon_tsdf* h, const float* host_d, const float* host_w,
                                 recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_upload: handle is NULL");
    const size_t bytes = sizeof(float) * (size_t)h->info.num_voxels;
    hipStream_t s = (hipStream_t)stream;
    if (host_d) RECON_HIP_TRY(hipMemcpyAsync(h->D, host_d, bytes, hipMemcpyHostToDevice, s), "upload D");
    if (host_w) RECON_HIP_TRY(hipMemcpyAsync(h->W, host_w, bytes, hipMemcpyHostToDevice, s), "upload W");
    RECON_HIP_TRY(hipStreamSynchronize(s), "recon_tsdf_upload: synchronize");
    return RECON_OK;
}

extern "C" int recon_tsdf_set_touch_counter(recon_tsdf* h, unsigned long long* counter) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_set_touch_counter: handle is NULL");
    h->touch_counter = counter;
    return RECON_OK;
}

extern "C" int recon_tsdf_set_update_counter(recon_tsdf* h, unsigned long long* counter) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_set_update_counter: handle is NULL");
    h->update_counter = counter;
    return RECON_OK;
}

// TSDFVolume::integrate (tsdf_volume.cu:257-290): the reference rejects a
// null colour image; colour is never read, so here it is optional.
extern "C" int recon_tsdf_integrate(recon_tsdf* h, const recon_depth_view* depth,
                                    const uint8_t* rgb, size_t rgb_step,
                                    const double camera_pose[7], recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_integrate: handle is NULL");
    int rc = check_depth(depth, "recon_tsdf_integrate");
    if (rc) return rc;
    if (!camera_pose) return fail(RECON_ERR_NULL, "recon_tsdf_integrate: pose is NULL");
    if (!pose_is_finite(camera_pose))
        return fail(RECON_ERR_CONFIG, "recon_tsdf_integrate: pose has a non-finite entry");
    const bool color = h->C && rgb;  // the image is ignored unless colour is enabled
    if (color && (rgb_step < 3 * (size_t)depth->cols || rgb_step * (size_t)depth->rows >= (size_t(1) << 31)))
        return fail(RECON_ERR_SIZE, "recon_tsdf_integrate: rgb step %zu invalid", rgb_step);
    const VolumeParams vp = volume_params(h, depth->cols, depth->rows);
    FrameBlock<1> fb;
    const int pitch = (int32_t)(depth->step / 4);
    make_frame_params(vp, camera_pose, depth->data, pitch, &fb.f[0]);
    if (color) {
        fb.f[0].rgb = rgb;
        fb.f[0].rgb_pitch = (int32_t)rgb_step;
    }
    return launch_integrate<1>(h, vp, depth->data, 0, pitch, 1, fb, (hipStream_t)stream);
}

extern "C" int recon_tsdf_enable_color(recon_tsdf* h, int enable, recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_enable_color: handle is NULL");
    hipStream_t s = (hipStream_t)stream;
    if (!enable) {
        if (h->C) {
            RECON_HIP_TRY(hipStreamSynchronize(s), "colour release synchronize");
            (void)hipFree(h->C);
        }
        h->C = nullptr;
        h->vp.C = nullptr;
        return RECON_OK;
    }
    if (h->C) return RECON_OK;
    int rc = require_gfx950();
    if (rc) return rc;
    const long long n = h->info.num_voxels;
    RECON_HIP_TRY(hipMalloc((void**)&h->C, 16 * (size_t)n), "colour volume");
    h->vp.C = h->C;
    hipLaunchKernelGGL(fill_color_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<float4*>(h->C), n);
    RECON_HIP_TRY(hipGetLastError(), "fill_color_kernel launch");
    return RECON_OK;
}

extern "C" int recon_tsdf_transformation_voxels(recon_tsdf* h, void* out, recon_stream_t stream) {
    clear_error();
    if (!h || !out) return fail(RECON_ERR_NULL, "recon_tsdf_transformation_voxels: NULL argument");
    const long long n = h->info.num_voxels;
    hipLaunchKernelGGL(transformation_voxels_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, h->vp, static_cast<float*>(out), n);
    RECON_HIP_TRY(hipGetLastError(), "transformation_voxels_kernel launch");
    return RECON_OK;
}

extern "C" float* recon_tsdf_colors(recon_tsdf* h) { return h ? h->C : nullptr; }

extern "C" int recon_tsdf_download_colors(recon_tsdf* h, float* host, recon_stream_t stream) {
    clear_error();
    if (!h || !host) return fail(RECON_ERR_NULL, "recon_tsdf_download_colors: NULL argument");
    if (!h->C) return fail(RECON_ERR_CONFIG, "recon_tsdf_download_colors: colour is not enabled");
    hipStream_t s = (hipStream_t)stream;
    RECON_HIP_TRY(hipMemcpyAsync(host, h->C, 16 * (size_t)h->info.num_voxels, hipMemcpyDeviceToHost, s),
                  "download colours");
    RECON_HIP_TRY(hipStreamSynchronize(s), "recon_tsdf_download_colors: synchronize");
    return RECON_OK;
}

extern "C" int recon_tsdf_integrate_frames(recon_tsdf* h, const recon_depth_view* depth0,
                                           size_t frame_stride_bytes, int32_t num_frames,
                                           const double* camera_poses, recon_stream_t stream) {
    clear_error();
    if (!h) return fail(RECON_ERR_NULL, "recon_tsdf_integrate_frames: handle is NULL");
    int rc = check_depth(depth0, "recon_tsdf_integrate_frames");
    if (rc) return rc;
    if (!camera_poses) return fail(RECON_ERR_NULL, "recon_tsdf_integrate_frames: poses is NULL");
    if (num_frames < 0) return fail(RECON_ERR_SIZE, "recon_tsdf_integrate_frames: num_frames < 0");
    if (num_frames > 1 && frame_stride_bytes % 4 != 0)
        return fail(RECON_ERR_SIZE, "recon_tsdf_integrate_frames: frame stride not a multiple of 4");
    for (int f = 0; f < num_frames; ++f)
        if (!pose_is_finite(camera_poses + 7 * f))
            return fail(RECON_ERR_CONFIG, "recon_tsdf_integrate_frames: pose %d non-finite", f);
    const VolumeParams vp = volume_params(h, depth0->cols, depth0->rows);
    const char* base = reinterpret_cast<const char*>(depth0->data);
    for (int f0 = 0; f0 < num_frames; f0 += kMaxFusedFrames) {
        const int n = std::min(kMaxFusedFrames, num_frames - f0);
        FrameBlock<kMaxFusedFrames> fb;
        std::memset(&fb, 0, sizeof(fb));
        for (int k = 0; k < n; ++k) {
            const float* dptr = reinterpret_cast<const float*>(base + (size_t)(f0 + k) * frame_stride_bytes);
            make_frame_params(vp, camera_poses + 7 * (f0 + k), dptr, (int32_t)(depth0->step / 4),
                              &fb.f[k]);
        }
        rc = launch_integrate<kMaxFusedFrames>(
            h, vp, reinterpret_cast<const float*>(base + (size_t)f0 * frame_stride_bytes),
            frame_stride_bytes / 4, (int32_t)(depth0->step / 4), n, fb, (hipStream_t)stream);
        if (rc) return rc;
    }
    return RECON_OK;
}
