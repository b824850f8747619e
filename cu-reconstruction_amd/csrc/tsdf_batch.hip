// tsdf_batch.hip -- many independent TSDF volumes of uneven sizes, frame
// counts and image resolutions, integrated and meshed by one hipGraph launch.
//
// The reference builds one TSDFVolume per scene and blocks on every frame
// (tsdf_surface_reconstruction.cpp:125-138, tsdf_volume.cu:287); a batch of
// scenes there is a loop over volumes and frames.  Here a batch is planned
// once (device tables of volume and frame parameters, work tables, one
// arena for every D and W) and captured as a graph whose replay runs:
//   fill        D <- truncation, W <- 0 for every volume (one launch);
//   bounds      depth bounds of every frame of every volume (one launch,
//               also clears the cull-list counters);
//   per tile-shape group: cull (frustum + depth tests of every wave-tile of
//               the group's volumes against all frames of its volume) and a
//               persistent integrate over the list: each wave-tile loads D
//               and W once, applies all of its volume's frames in order and
//               stores once;
//   marching cubes over every volume (chunk counts per tile of every
//               volume, one scan across all tiles, emission), then each
//               volume's triangle range.
// Every voxel sees the same float32 operations, frame by frame, as the
// single-volume path, so D, W and meshes are identical to looping over
// volumes with recon_tsdf_integrate / recon_tsdf_extract_mesh.
//
// Tile shape (x lanes per row, z rows per lane) and the integrate grid's
// share of the resident wave slots are chosen per (grid size, image size)
// class by recon_batch_autotune, which times the variants and stores the
// winners in a JSON file that recon_batch_create reads.
#pragma clang fp contract(off)

#include "integrate_device.hpp"
#include "mc_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace recon {

struct BatchVolParams {
    VolumeParams vp;
    int32_t first_frame;
    int32_t nframes;
};

struct Variant {
    int lx, nr, occ_div;
};
// searched by the autotuner; index 0 is the default
constexpr Variant kVariants[] = {{8, 2, 1}, {16, 2, 1}, {8, 1, 1}, {16, 1, 1},
                                 {8, 2, 2}, {16, 2, 2}, {8, 1, 2}, {16, 1, 2}};
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
constexpr int kFillBlock = 1024;  // voxels per fill workgroup

// ---- kernels -------------------------------------------------------------------
__global__ __launch_bounds__(256) void batch_fill_kernel(const BatchVolParams* __restrict__ vols,
                                                         const int* __restrict__ block_start,
                                                         int nvol) {
    const int b = blockIdx.x;
    int lo = 0, hi = nvol - 1;  // last volume whose first block <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (block_start[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const VolumeParams& vp = vols[lo].vp;
    const long long n = (long long)vp.gx * vp.gy * vp.gz;
    fill4(vp.D, vp.W, (long long)(b - block_start[lo]) * kFillBlock + 4 * threadIdx.x, n, vp.trunc);
}

// One workgroup per 32x32 block of every frame: job_start[f] is the first
// workgroup of frame f (frames of any size, no idle workgroups).
__global__ __launch_bounds__(256) void batch_bounds_kernel(const BoundsJob* __restrict__ jobs,
                                                           const int* __restrict__ job_start,
                                                           int njobs, unsigned* counters,
                                                           int ncounters) {
    __shared__ int s0[16];
    __shared__ int s1[4];
    if (blockIdx.x == 0 && (int)threadIdx.x < ncounters) counters[threadIdx.x] = 0;
    const int b = blockIdx.x;
    int lo = 0, hi = njobs - 1;  // last frame whose first workgroup <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (job_start[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const BoundsJob j = jobs[lo];
    const int k = b - job_start[lo];
    bounds_block(j, k % j.b1x, k / j.b1x, s0, s1);
}

template <class S>
__global__ __launch_bounds__(256) void batch_cull_kernel(const BatchVolParams* __restrict__ vols,
                                                         const FrameParams* __restrict__ frames,
                                                         const uint2* __restrict__ wg_table,
                                                         ListEntry* __restrict__ list,
                                                         unsigned* __restrict__ list_count) {
    __shared__ unsigned s_wave[kThreads / 64];
    __shared__ unsigned s_base;
    const uint2 e = wg_table[blockIdx.x];  // (volume, first tile)
    const BatchVolParams& bv = vols[e.x];
    const uint32_t ntiles = (uint32_t)bv.vp.tiles_x * bv.vp.tiles_y * bv.vp.tiles_z;
    const uint32_t id = e.y + threadIdx.x;
    const uint64_t live =
        id < ntiles ? tile_frames<S>(bv.vp, bv.nframes, FrameTable{frames + bv.first_frame}, id) : 0ull;
    append_tile(live, id, e.x, list, list_count, s_wave, &s_base);
}

template <class S>
__global__ __launch_bounds__(kThreads) void batch_integrate_kernel(
    const BatchVolParams* __restrict__ vols, const FrameParams* __restrict__ frames,
    const ListEntry* __restrict__ list, const unsigned* __restrict__ list_count) {
    const unsigned n = __builtin_amdgcn_readfirstlane(*list_count);
    const unsigned waves = gridDim.x * (kThreads / 64);
    unsigned cnt = 0;
    for (unsigned i = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t tile = __builtin_amdgcn_readfirstlane(list[i].tile);
        const uint32_t v = __builtin_amdgcn_readfirstlane(list[i].volume);
        const uint64_t mask =
            ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(list[i].frames >> 32)) << 32) |
            __builtin_amdgcn_readfirstlane((uint32_t)list[i].frames);
        const BatchVolParams& bv = vols[v];
        int x0, y0, z0;
        tile_origin<S>(bv.vp, tile, x0, y0, z0);
        const FrameTable ft{frames + bv.first_frame};
        if (bv.vp.gx % 4 == 0)
            integrate_tile<S, FrameTable, true, false>(bv.vp, bv.nframes, ft, mask, x0, y0, z0, cnt,
                                                       cnt);
        else
            integrate_tile<S, FrameTable, false, false>(bv.vp, bv.nframes, ft, mask, x0, y0, z0, cnt,
                                                        cnt);
    }
}

// Marching cubes of every volume: one 256-thread workgroup per tile of a
// volume (mc_count_tile / mc_emit_tile, the single-volume path's code), the
// tiles of all volumes in one table and one scan, so a volume's triangles
// are the range of its first tile's offset to its last tile's end.
struct McBatchVol {
    McParams p;
    McTiling t;
    long long cc_base;  // first chunk count of the volume
};

__global__ __launch_bounds__(256) void batch_mc_count_kernel(const McBatchVol* __restrict__ vols,
                                                             const int2* __restrict__ tiles,
                                                             uint16_t* __restrict__ cc,
                                                             int32_t* __restrict__ tile_sum) {
    __shared__ uint8_t ntab[256];
    __shared__ int wtot[8];
    ntab[threadIdx.x] = mc_ntab_qbyte((int)threadIdx.x);
    __syncthreads();
    const int2 e = tiles[blockIdx.x];  // (volume, tile of the volume)
    const McBatchVol& v = vols[e.x];
    mc_count_tile(v.p, v.t, e.y, cc + v.cc_base, ntab, wtot, tile_sum + blockIdx.x, nullptr);
}

__global__ __launch_bounds__(256) void batch_mc_emit_kernel(const McBatchVol* __restrict__ vols,
                                                            const int2* __restrict__ tiles,
                                                            const uint16_t* __restrict__ cc,
                                                            const long long* __restrict__ tile_off,
                                                            float* __restrict__ verts,
                                                            long long capacity) {
    __shared__ McEmitLds lds;
    const long long base = tile_off[blockIdx.x];
    if (tile_off[blockIdx.x + 1] == base) return;  // the same for every thread
    mc_emit_lds_init(lds);
    const int2 e = tiles[blockIdx.x];
    const McBatchVol& v = vols[e.x];
    mc_emit_tile(v.p, v.t, e.y, cc + v.cc_base, base, lds, verts, nullptr, capacity);
}

// ranges[2v], ranges[2v+1] = first triangle and triangle count of volume v
// (its tiles are first_tile[v] .. first_tile[v+1]-1).
__global__ void batch_mesh_ranges_kernel(const int* __restrict__ first_tile, int nvol,
                                         const long long* __restrict__ tile_off,
                                         long long* __restrict__ ranges) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvol) return;
    const long long a = tile_off[first_tile[v]];
    const long long b = tile_off[first_tile[v + 1]];
    ranges[2 * v] = a;
    ranges[2 * v + 1] = b - a;
}

// ---- planning ---------------------------------------------------------------------
template <int LX, int NR>
static const void* cull_fn() { return reinterpret_cast<const void*>(&batch_cull_kernel<TileShape<LX, NR>>); }
template <int LX, int NR>
static const void* integrate_fn() {
    return reinterpret_cast<const void*>(&batch_integrate_kernel<TileShape<LX, NR>>);
}

static int variant_shape(int v, int& X, int& Y, int& Z) {
    X = 4 * kVariants[v].lx;
    Y = 64 / kVariants[v].lx;
    Z = kVariants[v].nr;
    return 0;
}

template <int LX, int NR>
static int launch_group(const BatchVolParams* vols, const FrameParams* frames, const uint2* wg,
                        int nwg, ListEntry* list, unsigned* count, int grid, hipStream_t s) {
    using S = TileShape<LX, NR>;
    hipLaunchKernelGGL((batch_cull_kernel<S>), dim3(nwg), dim3(256), 0, s, vols, frames, wg, list,
                       count);
    RECON_HIP_TRY(hipGetLastError(), "batch_cull_kernel launch");
    hipLaunchKernelGGL((batch_integrate_kernel<S>), dim3(grid), dim3(kThreads), 0, s, vols, frames,
                       list, count);
    RECON_HIP_TRY(hipGetLastError(), "batch_integrate_kernel launch");
    return RECON_OK;
}

static int group_occupancy(int v) {
    const void* fn = nullptr;
    switch (kVariants[v].lx * 10 + kVariants[v].nr) {
        case 82: fn = integrate_fn<8, 2>(); break;
        case 162: fn = integrate_fn<16, 2>(); break;
        case 81: fn = integrate_fn<8, 1>(); break;
        default: fn = integrate_fn<16, 1>(); break;
    }
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, kThreads, 0) != hipSuccess ||
        blocks < 1)
        blocks = 4;
    return std::max(1, blocks / kVariants[v].occ_div);
}

static int dispatch_group(int v, const BatchVolParams* vols, const FrameParams* frames,
                          const uint2* wg, int nwg, ListEntry* list, unsigned* count, int grid,
                          hipStream_t s) {
    switch (kVariants[v].lx * 10 + kVariants[v].nr) {
        case 82: return launch_group<8, 2>(vols, frames, wg, nwg, list, count, grid, s);
        case 162: return launch_group<16, 2>(vols, frames, wg, nwg, list, count, grid, s);
        case 81: return launch_group<8, 1>(vols, frames, wg, nwg, list, count, grid, s);
        default: return launch_group<16, 1>(vols, frames, wg, nwg, list, count, grid, s);
    }
}

static std::string class_key(const recon_tsdf_info& info, int cols, int rows) {
    char buf[96];
    std::snprintf(buf, sizeof(buf), "g%dx%dx%d_%dx%d", info.grid[0], info.grid[1], info.grid[2],
                  cols, rows);
    return buf;
}

// Reads {"<key>": {"variant": N, ...}, ...} as written by save_cache.
static std::map<std::string, int> load_cache(const char* path) {
    std::map<std::string, int> m;
    if (!path) return m;
    std::ifstream in(path);
    if (!in) return m;
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string t = ss.str();
    size_t pos = 0;
    while ((pos = t.find("\"g", pos)) != std::string::npos) {
        const size_t end = t.find('"', pos + 1);
        if (end == std::string::npos) break;
        const std::string key = t.substr(pos + 1, end - pos - 1);
        const size_t vp = t.find("\"variant\"", end);
        const size_t next = t.find("\"g", end);
        if (vp != std::string::npos && (next == std::string::npos || vp < next)) {
            const size_t colon = t.find(':', vp);
            const int v = std::atoi(t.c_str() + colon + 1);
            if (v >= 0 && v < kNumVariants) m[key] = v;
        }
        pos = end + 1;
    }
    return m;
}

}  // namespace recon

using namespace recon;

struct recon_batch {
    int n = 0;
    bool with_mesh = true;
    std::vector<recon_tsdf_info> info;
    std::vector<int> variant;
    std::vector<void*> allocs;
    float* arena = nullptr;
    BatchVolParams* d_vols = nullptr;
    std::vector<BatchVolParams> h_vols;
    FrameParams* d_frames = nullptr;
    BoundsJob* d_jobs = nullptr;
    int* d_job_start = nullptr;  // first bounds workgroup of each frame
    int total_frames = 0, bounds_blocks = 0;
    int* d_fill_start = nullptr;
    int fill_blocks = 0;
    struct Group {
        int variant;
        uint2* wg;
        int nwg;
        ListEntry* list;
        int grid;
    };
    std::vector<Group> groups;
    unsigned* d_counts = nullptr;
    McBatchVol* d_mc = nullptr;
    int2* d_mc_tiles = nullptr;
    int mc_tiles = 0;
    uint16_t* mc_cc = nullptr;
    int32_t* tile_sum = nullptr;
    long long* tile_off = nullptr;
    int* d_first_tile = nullptr;
    long long* d_ranges = nullptr;
    float* verts = nullptr;
    long long capacity = 0;
    std::vector<long long> ranges;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~recon_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : allocs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** p, size_t count, const char* what) {
        void* q = nullptr;
        RECON_HIP_TRY(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)), what);
        allocs.push_back(q);
        *p = static_cast<T*>(q);
        return RECON_OK;
    }
    template <class T>
    int upload(T** p, const std::vector<T>& v, const char* what) {
        int rc = alloc(p, v.size(), what);
        if (rc) return rc;
        if (!v.empty())
            RECON_HIP_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), what);
        return RECON_OK;
    }
    int record(hipStream_t s);
};

int recon_batch::record(hipStream_t s) {
    hipLaunchKernelGGL(batch_fill_kernel, dim3(fill_blocks), dim3(256), 0, s, d_vols, d_fill_start, n);
    RECON_HIP_TRY(hipGetLastError(), "batch_fill_kernel launch");
    if (total_frames > 0) {
        hipLaunchKernelGGL(batch_bounds_kernel, dim3(bounds_blocks), dim3(256), 0, s, d_jobs,
                           d_job_start, total_frames, d_counts, (int)groups.size());
        RECON_HIP_TRY(hipGetLastError(), "batch_bounds_kernel launch");
        for (size_t g = 0; g < groups.size(); ++g) {
            const Group& G = groups[g];
            if (G.nwg == 0) continue;
            int rc = dispatch_group(G.variant, d_vols, d_frames, G.wg, G.nwg, G.list, d_counts + g,
                                    G.grid, s);
            if (rc) return rc;
        }
    }
    if (with_mesh) {
        if (mc_tiles > 0) {
            hipLaunchKernelGGL(batch_mc_count_kernel, dim3(mc_tiles), dim3(256), 0, s, d_mc,
                               d_mc_tiles, mc_cc, tile_sum);
            RECON_HIP_TRY(hipGetLastError(), "batch_mc_count_kernel launch");
        }
        int rc = launch_tile_scan(tile_sum, mc_tiles, tile_off, s);
        if (rc) return rc;
        if (mc_tiles > 0) {
            hipLaunchKernelGGL(batch_mc_emit_kernel, dim3(mc_tiles), dim3(256), 0, s, d_mc,
                               d_mc_tiles, mc_cc, tile_off, verts, capacity);
            RECON_HIP_TRY(hipGetLastError(), "batch_mc_emit_kernel launch");
        }
        hipLaunchKernelGGL(batch_mesh_ranges_kernel, dim3((n + 63) / 64), dim3(64), 0, s,
                           d_first_tile, n, tile_off, d_ranges);
        RECON_HIP_TRY(hipGetLastError(), "batch_mesh_ranges_kernel launch");
    }
    return RECON_OK;
}

static int check_volume(const recon_batch_volume& v, int i, recon_tsdf_info* info) {
    int rc = describe(&v.config, info);
    if (rc) return fail(rc, "volume %d: %s", i, recon_last_error());
    if (info->weight_type != RECON_WEIGHT_BY_DEPTH)
        return fail(RECON_ERR_CONFIG, "volume %d: batches integrate with WEIGHT_BY_DEPTH only", i);
    if (v.num_frames < 0 || v.num_frames > 64)
        return fail(RECON_ERR_SIZE, "volume %d: num_frames %d outside 0..64", i, v.num_frames);
    if (v.num_frames > 0) {
        if (!v.depth || !v.poses) return fail(RECON_ERR_NULL, "volume %d: depth or poses is NULL", i);
        if (v.rows <= 0 || v.cols <= 0) return fail(RECON_ERR_SIZE, "volume %d: bad image size", i);
        if (v.depth_step < 4 * (size_t)v.cols || v.depth_step % 4 || v.frame_stride % 4)
            return fail(RECON_ERR_SIZE, "volume %d: bad depth step / frame stride", i);
        if ((v.depth_step / 4) * (size_t)v.rows >= (size_t(1) << 31))
            return fail(RECON_ERR_RANGE, "volume %d: depth image exceeds 2^31 floats", i);
        for (int f = 0; f < v.num_frames; ++f)
            if (!pose_is_finite(v.poses + 7 * f))
                return fail(RECON_ERR_CONFIG, "volume %d: pose %d non-finite", i, f);
    }
    return RECON_OK;
}

static int plan_batch(const recon_batch_volume* vols, int n, const std::vector<int>& variant,
                      bool with_mesh, int64_t capacity, recon_batch** out) {
    int num_cus = 0, dev = 0;
    RECON_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    RECON_HIP_TRY(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev),
                  "hipDeviceGetAttribute");
    recon_batch* b = new recon_batch();
    auto fail_out = [&](int rc) {
        delete b;
        return rc;
    };
    b->n = n;
    b->with_mesh = with_mesh;
    b->variant = variant;
    b->info.resize(n);
    for (int i = 0; i < n; ++i) {
        int rc = check_volume(vols[i], i, &b->info[i]);
        if (rc) return fail_out(rc);
    }
    // one arena: D of every volume, then W of every volume, each padded to
    // whole fill blocks
    std::vector<long long> vox_off(n + 1, 0);
    std::vector<int> fill_start(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const long long nv = b->info[i].num_voxels;
        const long long blocks = (nv + kFillBlock - 1) / kFillBlock;
        vox_off[i + 1] = vox_off[i] + blocks * kFillBlock;
        fill_start[i + 1] = fill_start[i] + (int)blocks;
    }
    b->fill_blocks = fill_start[n];
    int rc = b->alloc(&b->arena, 2 * (size_t)vox_off[n], "batch volumes");
    if (rc) return fail_out(rc);
    // volume and frame tables
    std::vector<FrameParams> frames;
    std::vector<BoundsJob> jobs;
    std::vector<int> job_start;
    size_t bounds_floats = 0;
    for (int i = 0; i < n; ++i) {
        const recon_tsdf_info& in = b->info[i];
        BatchVolParams bv;
        std::memset(&bv, 0, sizeof(bv));
        VolumeParams& vp = bv.vp;
        vp.D = b->arena + vox_off[i];
        vp.W = b->arena + vox_off[n] + vox_off[i];
        vp.gx = in.grid[0];
        vp.gy = in.grid[1];
        vp.gz = in.grid[2];
        int X, Y, Z;
        variant_shape(variant[i], X, Y, Z);
        vp.tiles_x = (vp.gx + X - 1) / X;
        vp.tiles_y = (vp.gy + Y - 1) / Y;
        vp.tiles_z = (vp.gz + Z - 1) / Z;
        for (int k = 0; k < 3; ++k) {
            vp.vs[k] = in.voxel_size[k];
            vp.off[k] = in.global_offset[k];
        }
        vp.fx = in.intrinsics[0];
        vp.fy = in.intrinsics[1];
        vp.cx = in.intrinsics[2];
        vp.cy = in.intrinsics[3];
        vp.trunc = in.truncation;
        vp.maxw = in.max_weight;
        vp.maxd = in.max_depth;
        const int cols = vols[i].num_frames ? vols[i].cols : 1, rows = vols[i].num_frames ? vols[i].rows : 1;
        vp.cols = cols;
        vp.rows = rows;
        vp.fcols = (float)cols;
        vp.frows = (float)rows;
        vp.b0x = (cols + kBound0 - 1) / kBound0;
        vp.b0y = (rows + kBound0 - 1) / kBound0;
        vp.b1x = (cols + kBound1 - 1) / kBound1;
        vp.b1y = (rows + kBound1 - 1) / kBound1;
        bv.first_frame = (int)frames.size();
        bv.nframes = vols[i].num_frames;
        for (int f = 0; f < vols[i].num_frames; ++f) {
            const float* d = reinterpret_cast<const float*>(reinterpret_cast<const char*>(vols[i].depth) +
                                                            (size_t)f * vols[i].frame_stride);
            FrameParams fp;
            make_frame_params(vp, vols[i].poses + 7 * f, d, (int32_t)(vols[i].depth_step / 4), &fp);
            // bound pointers are filled once the arena exists (offset for now)
            fp.bound0 = reinterpret_cast<const float*>(bounds_floats);
            frames.push_back(fp);
            BoundsJob j;
            j.depth = d;
            j.bound0 = reinterpret_cast<float*>(bounds_floats);
            j.pitch = (int32_t)(vols[i].depth_step / 4);
            j.cols = cols;
            j.rows = rows;
            j.b0x = vp.b0x;
            j.b0y = vp.b0y;
            j.b1x = vp.b1x;
            j.b1y = vp.b1y;
            j.maxd = vp.maxd;
            j.vec = (reinterpret_cast<uintptr_t>(d) % 16 == 0 && j.pitch % 4 == 0) ? 1 : 0;
            jobs.push_back(j);
            job_start.push_back(b->bounds_blocks);
            b->bounds_blocks += vp.b1x * vp.b1y;
            bounds_floats += (size_t)vp.b0x * vp.b0y + (size_t)vp.b1x * vp.b1y;
        }
        b->h_vols.push_back(bv);
    }
    b->total_frames = (int)frames.size();
    float* bounds = nullptr;
    rc = b->alloc(&bounds, bounds_floats, "batch depth bounds");
    if (rc) return fail_out(rc);
    for (size_t f = 0; f < frames.size(); ++f) {
        const size_t o = reinterpret_cast<size_t>(frames[f].bound0);
        const int vi = [&] {
            int v = 0;
            while (v + 1 < n && b->h_vols[v + 1].first_frame <= (int)f) ++v;
            return v;
        }();
        const VolumeParams& vp = b->h_vols[vi].vp;
        frames[f].bound0 = bounds + o;
        frames[f].bound1 = bounds + o + (size_t)vp.b0x * vp.b0y;
        jobs[f].bound0 = bounds + o;
        jobs[f].bound1 = bounds + o + (size_t)vp.b0x * vp.b0y;
    }
    if ((rc = b->upload(&b->d_vols, b->h_vols, "batch volume table")) ||
        (rc = b->upload(&b->d_frames, frames, "batch frame table")) ||
        (rc = b->upload(&b->d_jobs, jobs, "batch bounds jobs")) ||
        (rc = b->upload(&b->d_job_start, job_start, "batch bounds job starts")) ||
        (rc = b->upload(&b->d_fill_start, fill_start, "batch fill table")))
        return fail_out(rc);
    // cull groups: one per variant present
    for (int v = 0; v < kNumVariants; ++v) {
        std::vector<uint2> wg;
        size_t ntiles_total = 0;
        for (int i = 0; i < n; ++i) {
            if (variant[i] != v || vols[i].num_frames == 0) continue;
            const VolumeParams& vp = b->h_vols[i].vp;
            const uint32_t nt = (uint32_t)vp.tiles_x * vp.tiles_y * vp.tiles_z;
            for (uint32_t t0 = 0; t0 < nt; t0 += 256) wg.push_back(make_uint2((uint32_t)i, t0));
            ntiles_total += nt;
        }
        if (wg.empty()) continue;
        recon_batch::Group G;
        G.variant = v;
        G.nwg = (int)wg.size();
        if ((rc = b->upload(&G.wg, wg, "batch cull table")) ||
            (rc = b->alloc(&G.list, ntiles_total, "batch cull list")))
            return fail_out(rc);
        G.grid = num_cus * group_occupancy(v);
        b->groups.push_back(G);
    }
    if ((rc = b->alloc(&b->d_counts, std::max<size_t>(b->groups.size(), 1), "batch list counters")))
        return fail_out(rc);
    // marching cubes tables: every volume's tiles, its chunk counts' place
    if (with_mesh) {
        std::vector<McBatchVol> mcs(n);
        std::vector<int2> tiles;
        std::vector<int> first_tile(n + 1, 0);
        long long cc_total = 0;
        for (int i = 0; i < n; ++i) {
            const VolumeParams& vp = b->h_vols[i].vp;
            McParams& p = mcs[i].p;
            p.f = vp.D;
            p.gx = vp.gx;
            p.gy = vp.gy;
            p.gz = vp.gz;
            p.cx = vp.gx - 1;
            p.cyr = vp.gy - 1;
            p.rows = (vp.gx >= 2 && vp.gy >= 2 && vp.gz >= 2) ? (vp.gy - 1) * (vp.gz - 1) : 0;
            for (int k = 0; k < 3; ++k) {
                p.vs[k] = vp.vs[k];
                p.off[k] = vp.off[k];
            }
            mcs[i].t = mc_tiling(vp.gx, vp.gy, vp.gz);
            mcs[i].cc_base = cc_total;
            cc_total += (long long)p.rows * mcs[i].t.nch;
            first_tile[i] = (int)tiles.size();
            for (int t = 0; t < mcs[i].t.tiles; ++t) tiles.push_back(make_int2(i, t));
        }
        first_tile[n] = (int)tiles.size();
        if (tiles.size() >= (size_t(1) << 31) - 1)
            return fail_out(fail(RECON_ERR_RANGE, "recon_batch_create: too many marching-cubes tiles"));
        b->mc_tiles = (int)tiles.size();
        b->capacity = capacity;
        if ((rc = b->upload(&b->d_mc, mcs, "batch mc table")) ||
            (rc = b->upload(&b->d_mc_tiles, tiles, "batch mc tiles")) ||
            (rc = b->upload(&b->d_first_tile, first_tile, "batch mc first tiles")) ||
            (rc = b->alloc(&b->mc_cc, (size_t)cc_total, "batch chunk counts")) ||
            (rc = b->alloc(&b->tile_sum, (size_t)b->mc_tiles, "batch tile sums")) ||
            (rc = b->alloc(&b->tile_off, (size_t)b->mc_tiles + 1, "batch tile offsets")) ||
            (rc = b->alloc(&b->d_ranges, 2 * (size_t)n, "batch mesh ranges")) ||
            (rc = b->alloc(&b->verts, 9 * (size_t)std::max<int64_t>(capacity, 1), "batch mesh")))
            return fail_out(rc);
    }
    // capture the whole batch as one graph
    hipStream_t s = nullptr;
    RECON_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "capture stream");
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = b->record(s);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s, &g);
        b->graph = g;
        if (rc == RECON_OK && e2 != hipSuccess) rc = hip_fail(e2, "hipStreamEndCapture");
        if (rc == RECON_OK) {
            const hipError_t e3 = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
            if (e3 != hipSuccess) rc = hip_fail(e3, "hipGraphInstantiate");
        }
    } else {
        rc = hip_fail(e, "hipStreamBeginCapture");
    }
    (void)hipStreamDestroy(s);
    if (rc) return fail_out(rc);
    b->ranges.assign(2 * (size_t)n, 0);
    *out = b;
    return RECON_OK;
}

static int batch_variants(const recon_batch_volume* vols, int n, const char* cache,
                          std::vector<int>& variant) {
    const std::map<std::string, int> m = load_cache(cache);
    variant.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        recon_tsdf_info info;
        int rc = describe(&vols[i].config, &info);
        if (rc) return fail(rc, "volume %d: %s", i, recon_last_error());
        auto it = m.find(class_key(info, vols[i].cols, vols[i].rows));
        if (it != m.end()) variant[i] = it->second;
    }
    return RECON_OK;
}

extern "C" int recon_batch_create(const recon_batch_volume* volumes, int32_t num_volumes,
                                  int64_t mesh_capacity_triangles, const char* tune_cache_path,
                                  recon_batch** out) {
    clear_error();
    if (!out) return fail(RECON_ERR_NULL, "recon_batch_create: out is NULL");
    *out = nullptr;
    if (!volumes || num_volumes <= 0) return fail(RECON_ERR_NULL, "recon_batch_create: no volumes");
    if (mesh_capacity_triangles < 0)
        return fail(RECON_ERR_SIZE, "recon_batch_create: negative mesh capacity");
    std::vector<int> variant;
    int rc = batch_variants(volumes, num_volumes, tune_cache_path, variant);
    if (rc) return rc;
    for (int i = 0; i < num_volumes; ++i) {
        recon_tsdf_info info;
        rc = check_volume(volumes[i], i, &info);
        if (rc) return rc;
    }
    rc = require_gfx950();
    if (rc) return rc;
    return plan_batch(volumes, num_volumes, variant, mesh_capacity_triangles > 0,
                      mesh_capacity_triangles, out);
}

extern "C" void recon_batch_destroy(recon_batch* b) { delete b; }

extern "C" int recon_batch_run(recon_batch* b, recon_stream_t stream) {
    clear_error();
    if (!b) return fail(RECON_ERR_NULL, "recon_batch_run: batch is NULL");
    RECON_HIP_TRY(hipGraphLaunch(b->exec, (hipStream_t)stream), "hipGraphLaunch");
    return RECON_OK;
}

extern "C" int recon_batch_volume_data(recon_batch* b, int32_t i, float** distances,
                                       float** weights, recon_tsdf_info* info) {
    clear_error();
    if (!b) return fail(RECON_ERR_NULL, "recon_batch_volume_data: batch is NULL");
    if (i < 0 || i >= b->n) return fail(RECON_ERR_SIZE, "recon_batch_volume_data: index %d", i);
    if (distances) *distances = b->h_vols[i].vp.D;
    if (weights) *weights = b->h_vols[i].vp.W;
    if (info) *info = b->info[i];
    return RECON_OK;
}

extern "C" int recon_batch_meshes(recon_batch* b, int64_t* ranges, float** verts,
                                  recon_stream_t stream) {
    clear_error();
    if (!b || !ranges) return fail(RECON_ERR_NULL, "recon_batch_meshes: null argument");
    if (!b->with_mesh) return fail(RECON_ERR_SIZE, "recon_batch_meshes: batch built without meshes");
    hipStream_t s = (hipStream_t)stream;
    RECON_HIP_TRY(hipMemcpyAsync(ranges, b->d_ranges, 2 * sizeof(long long) * (size_t)b->n,
                                 hipMemcpyDeviceToHost, s),
                  "mesh ranges");
    RECON_HIP_TRY(hipStreamSynchronize(s), "recon_batch_meshes synchronize");
    if (verts) *verts = b->verts;
    long long total = 0;
    for (int i = 0; i < b->n; ++i) total = std::max<long long>(total, ranges[2 * i] + ranges[2 * i + 1]);
    if (total > b->capacity)
        return fail(RECON_ERR_CAPACITY, "recon_batch_meshes: %lld triangles exceed capacity %lld",
                    total, b->capacity);
    return RECON_OK;
}

// Whole-graph time of `vols` with tile-shape `variant` per volume
// (integrate only: fill, bounds, cull and integrate launches).
static int time_plan(const recon_batch_volume* vols, int n, const std::vector<int>& variant,
                     int iterations, hipStream_t s, hipEvent_t t0, hipEvent_t t1, float* us) {
    recon_batch* b = nullptr;
    int rc = plan_batch(vols, n, variant, false, 0, &b);
    if (rc) return rc;
    rc = recon_batch_run(b, s);  // warm-up
    if (rc == RECON_OK && hipEventRecord(t0, s) != hipSuccess) rc = fail(RECON_ERR_HIP, "event");
    for (int it = 0; rc == RECON_OK && it < iterations; ++it) rc = recon_batch_run(b, s);
    if (rc == RECON_OK && hipEventRecord(t1, s) != hipSuccess) rc = fail(RECON_ERR_HIP, "event");
    if (rc == RECON_OK && hipEventSynchronize(t1) != hipSuccess) rc = fail(RECON_ERR_HIP, "event");
    float ms = 0.f;
    if (rc == RECON_OK) (void)hipEventElapsedTime(&ms, t0, t1);
    delete b;
    *us = 1000.0f * ms / iterations;
    return rc;
}

// Greedy search in the context of the whole batch: starting from the cached
// (or default) variant of every class, each class in turn tries every
// variant and keeps one only if the whole graph gets at least 2 % faster.
// Mixing variants adds launches (one cull + integrate pair per variant), so
// a variant that wins on its class alone can lose in the batch; timing the
// whole batch accounts for that.
extern "C" int recon_batch_autotune(const recon_batch_volume* volumes, int32_t num_volumes,
                                    const char* cache_path, int32_t iterations,
                                    recon_stream_t stream) {
    clear_error();
    if (!volumes || num_volumes <= 0 || !cache_path)
        return fail(RECON_ERR_NULL, "recon_batch_autotune: null argument");
    int rc = require_gfx950();
    if (rc) return rc;
    iterations = std::max(1, iterations);
    hipStream_t s = (hipStream_t)stream;
    std::vector<int> variant;
    rc = batch_variants(volumes, num_volumes, cache_path, variant);
    if (rc) return rc;
    // classes of (grid, image) present, largest work first
    std::map<std::string, std::vector<int>> classes;
    std::map<std::string, double> work;
    for (int i = 0; i < num_volumes; ++i) {
        recon_tsdf_info info;
        rc = check_volume(volumes[i], i, &info);
        if (rc) return rc;
        if (volumes[i].num_frames == 0) continue;
        const std::string k = class_key(info, volumes[i].cols, volumes[i].rows);
        classes[k].push_back(i);
        work[k] += (double)info.num_voxels * volumes[i].num_frames;
    }
    std::vector<std::string> order;
    for (const auto& kv : classes) order.push_back(kv.first);
    std::sort(order.begin(), order.end(), [&](const std::string& a, const std::string& b) {
        return work[a] > work[b];
    });
    hipEvent_t t0, t1;
    RECON_HIP_TRY(hipEventCreate(&t0), "hipEventCreate");
    RECON_HIP_TRY(hipEventCreate(&t1), "hipEventCreate");
    float best_us = 0.f;
    rc = time_plan(volumes, num_volumes, variant, iterations, s, t0, t1, &best_us);
    std::map<std::string, std::vector<float>> times;
    std::map<std::string, int> chosen;
    for (size_t c = 0; rc == RECON_OK && c < order.size(); ++c) {
        const std::vector<int>& members = classes[order[c]];
        const int current = variant[members[0]];
        std::vector<float> us(kNumVariants, best_us);
        int pick = current;
        for (int v = 0; v < kNumVariants && rc == RECON_OK; ++v) {
            if (v == current) continue;
            std::vector<int> trial = variant;
            for (int i : members) trial[i] = v;
            rc = time_plan(volumes, num_volumes, trial, iterations, s, t0, t1, &us[v]);
            if (rc == RECON_OK && us[v] < 0.98f * best_us && us[v] < us[pick]) pick = v;
        }
        if (rc) break;
        if (pick != current) {
            for (int i : members) variant[i] = pick;
            best_us = us[pick];
        }
        times[order[c]] = us;
        chosen[order[c]] = pick;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    std::map<std::string, int> all = load_cache(cache_path);
    for (const auto& kv : chosen) all[kv.first] = kv.second;
    std::ofstream o(cache_path);
    if (!o) return fail(RECON_ERR_SIZE, "recon_batch_autotune: cannot write %s", cache_path);
    o << "{\n";
    size_t k = 0;
    for (const auto& kv : all) {
        const Variant& V = kVariants[kv.second];
        o << "  \"" << kv.first << "\": {\"variant\": " << kv.second << ", \"lanes_x\": " << V.lx
          << ", \"rows_per_lane\": " << V.nr << ", \"occupancy_divisor\": " << V.occ_div;
        auto it = times.find(kv.first);
        if (it != times.end()) {
            o << ", \"batch_us\": [";
            for (size_t j = 0; j < it->second.size(); ++j) o << (j ? ", " : "") << it->second[j];
            o << "]";
        }
        o << "}" << (++k < all.size() ? "," : "") << "\n";
    }
    o << "}\n";
    return o ? RECON_OK : fail(RECON_ERR_SIZE, "recon_batch_autotune: write failed");
}

extern "C" int recon_batch_download(recon_batch* b, int32_t i, float* host_d, float* host_w,
                                    recon_stream_t stream) {
    clear_error();
    if (!b) return fail(RECON_ERR_NULL, "recon_batch_download: batch is NULL");
    if (i < 0 || i >= b->n) return fail(RECON_ERR_SIZE, "recon_batch_download: index %d", i);
    const size_t bytes = sizeof(float) * (size_t)b->info[i].num_voxels;
    hipStream_t s = (hipStream_t)stream;
    if (host_d)
        RECON_HIP_TRY(hipMemcpyAsync(host_d, b->h_vols[i].vp.D, bytes, hipMemcpyDeviceToHost, s),
                      "batch download D");
    if (host_w)
        RECON_HIP_TRY(hipMemcpyAsync(host_w, b->h_vols[i].vp.W, bytes, hipMemcpyDeviceToHost, s),
                      "batch download W");
    RECON_HIP_TRY(hipStreamSynchronize(s), "recon_batch_download synchronize");
    return RECON_OK;
}

extern "C" int recon_batch_download_mesh(recon_batch* b, int32_t i, float* host_verts,
                                         int64_t capacity_triangles, int64_t* num_triangles,
                                         recon_stream_t stream) {
    clear_error();
    if (!b || !num_triangles) return fail(RECON_ERR_NULL, "recon_batch_download_mesh: null argument");
    if (i < 0 || i >= b->n) return fail(RECON_ERR_SIZE, "recon_batch_download_mesh: index %d", i);
    std::vector<int64_t> ranges(2 * (size_t)b->n);
    float* verts = nullptr;
    const int rc = recon_batch_meshes(b, ranges.data(), &verts, stream);
    if (rc) return rc;
    *num_triangles = ranges[2 * i + 1];
    if (ranges[2 * i + 1] > capacity_triangles)
        return fail(RECON_ERR_CAPACITY, "recon_batch_download_mesh: %lld triangles exceed %lld",
                    (long long)ranges[2 * i + 1], (long long)capacity_triangles);
    if (ranges[2 * i + 1] > 0) {
        if (!host_verts) return fail(RECON_ERR_NULL, "recon_batch_download_mesh: verts is NULL");
        RECON_HIP_TRY(hipMemcpy(host_verts, verts + 9 * ranges[2 * i],
                                sizeof(float) * 9 * (size_t)ranges[2 * i + 1], hipMemcpyDeviceToHost),
                      "batch mesh download");
    }
    return RECON_OK;
}
