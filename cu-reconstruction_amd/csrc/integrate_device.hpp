// integrate_device.hpp -- device code of the projective integrate shared by
// the single-volume path (tsdf_volume.hip) and the batched path
// (tsdf_batch.hip): depth bounds, frustum/depth culling of wave-tiles and the
// per-voxel update of cu_tsdf_integrate (cu-Reconstruction
// modules/src/cuda/tsdf_volume.cu:144-190, camera_util.cu:7-51,
// se3.cuh:82-100) under the arithmetic contract of recon_common.hpp.
#pragma once

#include "recon_common.hpp"

namespace recon {

constexpr int kThreads = 256;
constexpr int kRcpTable = 256;  // entries of an integrate workgroup's 1/(w+1) table

// Wave-tile of 4*LX x (64/LX) x NR voxels.  Lane l owns voxels x..x+3
// (x = x0 + 4*(l % LX)) of row y = y0 + l / LX for z = z0 .. z0+NR-1.  With
// LX = 8 every row segment of a wave is one aligned 128-byte line of D and W
// when gx is a multiple of 32.
template <int LX, int NR>
struct TileShape {
    static constexpr int kLanesX = LX;
    static constexpr int kRows = NR;
    static constexpr int X = 4 * LX;
    static constexpr int Y = 64 / LX;
    static constexpr int Z = NR;
};
using DefaultShape = TileShape<8, 2>;

// Frame parameters by value (kernel arguments) or from a device table; both
// are read as `frames.f[i]`.
template <int MAXF>
struct FrameBlock {
    static constexpr int kMax = MAXF;
    FrameParams f[MAXF];
};
struct FrameTable {
    static constexpr int kMax = 0;  // any number
    const FrameParams* f;
};

struct ListEntry {
    uint32_t tile;     // wave-tile index x + tiles_x * (y + tiles_y * z)
    uint32_t volume;   // volume index (batched path), 0 otherwise
    uint64_t frames;   // bit f: frame f of the volume may update the tile
};

// Max over a voxel-centre box of n.c + k.
__device__ __forceinline__ float plane_max(const float* p, const float lo[3], const float hi[3]) {
    float g = p[3];
    g = g + fmaxf(p[0] * lo[0], p[0] * hi[0]);
    g = g + fmaxf(p[1] * lo[1], p[1] * hi[1]);
    g = g + fmaxf(p[2] * lo[2], p[2] * hi[2]);
    return g;
}

template <class S>
__device__ __forceinline__ void tile_origin(const VolumeParams& vp, uint32_t id, int& x0, int& y0,
                                            int& z0) {
    const uint32_t tx = id % (uint32_t)vp.tiles_x;
    const uint32_t tyz = id / (uint32_t)vp.tiles_x;
    x0 = (int)tx * S::X;
    y0 = (int)(tyz % (uint32_t)vp.tiles_y) * S::Y;
    z0 = (int)(tyz / (uint32_t)vp.tiles_y) * S::Z;
}

// ---- depth bounds -------------------------------------------------------------
// One 32x32 pixel block (workgroup of 256 threads): the largest valid depth
// (0 < d <= maxd) of each 8x8 sub-block into bound0 and of the block into
// bound1, 0 when none is valid.  Valid depths are positive floats, so their
// bit patterns order like ints.
struct BoundsJob {
    const float* depth;
    float* bound0;   // b0x * b0y floats
    float* bound1;   // b1x * b1y floats (bound0 + b0x*b0y)
    int32_t pitch, cols, rows;
    int32_t b0x, b0y, b1x, b1y;
    float maxd;
    int32_t vec = 0;  // depth and pitch 16-byte aligned: four pixels per load
};

__device__ __forceinline__ int bounds_valid(float d, float maxd) {
    return (d > 0.0f && d <= maxd) ? __float_as_int(d) : 0;
}

// Thread t reads pixels 4*(t%8) .. +3 of row t/8 of the block, so wave w
// holds pixel rows 8w .. 8w+7: one row of four 8x8 sub-blocks, sub-block
// (t%8)/2.  Maxima are taken across lanes (no LDS atomics): s0 receives the
// 16 sub-block maxima and s1 the 4 wave maxima.
__device__ __forceinline__ void bounds_block(const BoundsJob& j, int bx, int by, int* s0, int* s1) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r = t >> 3, c4 = (t & 7) * 4;
    const int y = by * kBound1 + r;
    const int xb = bx * kBound1 + c4;
    int m = 0;
    if (y < j.rows) {
        const float* row = j.depth + (size_t)y * j.pitch;
        if (j.vec && xb + 3 < j.cols) {
            const float4 d = *reinterpret_cast<const float4*>(row + xb);
            m = max(max(bounds_valid(d.x, j.maxd), bounds_valid(d.y, j.maxd)),
                    max(bounds_valid(d.z, j.maxd), bounds_valid(d.w, j.maxd)));
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (xb + k < j.cols) m = max(m, bounds_valid(row[xb + k], j.maxd));
        }
    }
    // lanes of one sub-block share lane bits 1-2
    m = max(m, __shfl_xor(m, 1, 64));
    m = max(m, __shfl_xor(m, 8, 64));
    m = max(m, __shfl_xor(m, 16, 64));
    m = max(m, __shfl_xor(m, 32, 64));
    if (lane < 8 && (lane & 1) == 0) s0[wave * 4 + (lane >> 1)] = m;
    m = max(m, __shfl_xor(m, 2, 64));
    m = max(m, __shfl_xor(m, 4, 64));
    if (lane == 0) s1[wave] = m;
    __syncthreads();
    if (t < 16) {
        const int ox = bx * 4 + (t & 3), oy = by * 4 + (t >> 2);
        if (ox < j.b0x && oy < j.b0y) j.bound0[oy * j.b0x + ox] = __int_as_float(s0[t]);
    }
    if (t == 0) j.bound1[by * j.b1x + bx] = __int_as_float(max(max(s1[0], s1[1]), max(s1[2], s1[3])));
}

// ---- culling ------------------------------------------------------------------
// True if some block of [bx0,bx1] x [by0,by1] of a bound level holds a depth
// >= need.  Four independent loads per step keep the reads in flight together.
__device__ __forceinline__ bool bound_rect_reaches(const float* __restrict__ lvl, int stride,
                                                   int bx0, int bx1, int by0, int by1,
                                                   float need) {
    for (int by = by0; by <= by1; ++by) {
        const float* row = lvl + by * stride;
        for (int bx = bx0; bx <= bx1; bx += 4) {
            // indices past bx1 repeat bx1, which leaves the maximum unchanged
            const float a = row[bx];
            const float b = row[min(bx + 1, bx1)];
            const float c = row[min(bx + 2, bx1)];
            const float d = row[min(bx + 3, bx1)];
            if (fmaxf(fmaxf(a, b), fmaxf(c, d)) >= need) return true;
        }
    }
    return false;
}

// Cull test of a voxel-centre box for one frame (one thread).  Returns
// false only when no voxel of the box can be updated by the frame.  The
// box's eight corners bound the camera z of its voxels.  When all lie in
// front of the camera (z > 1e-3), the hull of their projections bounds the
// voxels' pixel coordinates (two pixels of slack cover float32 rounding, the
// hardware reciprocal used here (<= 1 ulp) and the truncation to an index),
// and the box survives when that footprint meets the image and the depth
// bound of its blocks reaches the box's nearest z minus the truncation; this
// also settles the frustum (a footprint outside the image fails, and a box
// beyond max depth + truncation finds no bound that far).  Otherwise the
// footprint is unbounded and the box survives when it meets the six frustum
// half-spaces of F.plane.  A footprint of more than kMaxFine 8x8 blocks keeps
// the box after the 32x32 test: the fine scan is a chain of dependent loads,
// and on a cull launch that fits the device in one pass its longest chain is
// the kernel's duration.
template <int kMaxFine = 96>
__device__ __forceinline__ bool box_reaches(const VolumeParams& vp, const FrameParams& F,
                                            const float lo[3], const float hi[3]) {
    float zmin = 3.0e38f, umin = 3.0e38f, umax = -3.0e38f, vmin = 3.0e38f, vmax = -3.0e38f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float cx = (i & 1) ? hi[0] : lo[0];
        const float cy = (i & 2) ? hi[1] : lo[1];
        const float cz = (i & 4) ? hi[2] : lo[2];
        const float px = ((F.R[0] * cx + F.R[1] * cy) + F.R[2] * cz) + F.t[0];
        const float py = ((F.R[3] * cx + F.R[4] * cy) + F.R[5] * cz) + F.t[1];
        const float pz = ((F.R[6] * cx + F.R[7] * cy) + F.R[8] * cz) + F.t[2];
        const float rz = __builtin_amdgcn_rcpf(pz);
        const float u = (px * vp.fx) * rz + vp.cx, v = (py * vp.fy) * rz + vp.cy;
        zmin = fminf(zmin, pz);
        umin = fminf(umin, u); umax = fmaxf(umax, u);
        vmin = fminf(vmin, v); vmax = fmaxf(vmax, v);
    }
    // corners at or behind the camera: only the frustum planes bound the box
    if (!(zmin > 1e-3f)) {
        bool in = true;
#pragma unroll
        for (int k = 0; k < 6; ++k) in = in && (plane_max(F.plane[k], lo, hi) >= 0.0f);
        return in;
    }
    umin = fmaxf(umin - 2.0f, 0.0f); umax = fminf(umax + 2.0f, vp.fcols - 1.0f);
    vmin = fmaxf(vmin - 2.0f, 0.0f); vmax = fminf(vmax + 2.0f, vp.frows - 1.0f);
    if (!(umin <= umax && vmin <= vmax)) return false;  // footprint leaves the image
    const int iu0 = (int)umin, iu1 = (int)umax, iv0 = (int)vmin, iv1 = (int)vmax;
    const float margin = 1e-5f * (fabsf(zmin) + vp.trunc) + F.zmargin;
    const float need = zmin - vp.trunc - margin;
    // 32x32 blocks first (a superset of the footprint's 8x8 blocks): a coarse
    // maximum below `need` settles the test with a few loads
    if (!bound_rect_reaches(F.bound1, vp.b1x, iu0 / kBound1, iu1 / kBound1, iv0 / kBound1,
                            iv1 / kBound1, need))
        return false;
    const int bx0 = iu0 / kBound0, bx1 = iu1 / kBound0, by0 = iv0 / kBound0, by1 = iv1 / kBound0;
    if ((bx1 - bx0 + 1) * (by1 - by0 + 1) > kMaxFine) return true;
    return bound_rect_reaches(F.bound0, vp.b0x, bx0, bx1, by0, by1, need);
}

// Frames (bit f of the result) whose frustum and depth bound reach wave-tile
// `id` of a volume; only frames first, first + step, ... are tested.
template <class S, class Frames, int kMaxFine = 96>
__device__ __forceinline__ uint64_t tile_frames(const VolumeParams& vp, const int nframes,
                                                const Frames& frames, uint32_t id,
                                                const int first = 0, const int step = 1) {
    int x0, y0, z0;
    tile_origin<S>(vp, id, x0, y0, z0);
    float lo[3], hi[3];
    lo[0] = voxel_centre(x0, vp.vs[0], vp.off[0]);
    lo[1] = voxel_centre(y0, vp.vs[1], vp.off[1]);
    lo[2] = voxel_centre(z0, vp.vs[2], vp.off[2]);
    hi[0] = voxel_centre(min(x0 + S::X - 1, vp.gx - 1), vp.vs[0], vp.off[0]);
    hi[1] = voxel_centre(min(y0 + S::Y - 1, vp.gy - 1), vp.vs[1], vp.off[1]);
    hi[2] = voxel_centre(min(z0 + S::Z - 1, vp.gz - 1), vp.vs[2], vp.off[2]);
    uint64_t live = 0;
    for (int f = first; f < nframes; f += step)
        if (box_reaches<kMaxFine>(vp, frames.f[f], lo, hi)) live |= 1ull << f;
    return live;
}

// Appends the calling workgroup's surviving tiles (`live` != 0) to `list`
// with one atomic per workgroup.  Every thread of the workgroup must call.
__device__ __forceinline__ void append_tile(uint64_t live, uint32_t tile, uint32_t volume,
                                            ListEntry* __restrict__ list,
                                            unsigned* __restrict__ list_count, unsigned* s_wave,
                                            unsigned* s_base) {
    const unsigned n = live != 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    unsigned wave_base = 0, total = 0;
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i) {
        wave_base += i < wave ? s_wave[i] : 0u;
        total += s_wave[i];
    }
    if (threadIdx.x == 0) *s_base = total ? atomicAdd(list_count, total) : 0u;
    __syncthreads();
    if (n) list[*s_base + wave_base + incl - n] = ListEntry{tile, volume, live};
}

// Appends the tiles of one wavefront (lane l: `tile`, frame mask `live`)
// that some frame reaches to `list`, with one atomic for the wavefront.
// Every lane of the calling wavefront must call.
__device__ __forceinline__ void append_wave(uint64_t live, uint32_t tile, uint32_t volume,
                                            ListEntry* __restrict__ list,
                                            unsigned* __restrict__ list_count) {
    const int lane = threadIdx.x & 63;
    const uint64_t ball = __ballot(live != 0);
    if (!ball) return;
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(list_count, (unsigned)__popcll(ball));
    base = (unsigned)__shfl((int)base, 0, 64);
    if (live) list[base + (unsigned)__popcll(ball & ((1ull << lane) - 1ull))] = ListEntry{tile, volume, live};
}

// Most wavefronts per workgroup of the frame-parallel culls: with W =
// blockDim.x / 64 of them (cull_waves), wave g of a workgroup tests frames
// g, g + W, ... of the workgroup's 64 tiles.
constexpr int kCullWaves = 16;

// W for `groups` workgroups of 64 tiles and at most `frames` frames a tile:
// about 32 wavefronts per CU in all (8 per SIMD), at most one per frame and
// at most kCullWaves (more wavefronts per tile only lengthen the workgroups
// when the groups already fill the device).
inline int cull_waves(long long groups, int frames, int cus) {
    long long w = groups > 0 ? (32LL * cus + groups - 1) / groups : 1;
    if (w > frames) w = frames;
    if (w > kCullWaves) w = kCullWaves;
    return w < 1 ? 1 : (int)w;
}

// ---- segmented cull list (no global counter) -------------------------------------
// A cull workgroup owns a segment of the list (the entries of its own tiles,
// at most one per tile) and writes only its count: no returning atomic, whose
// round trip would sit on every workgroup's critical path.  The consumer
// turns the counts into segment offsets in LDS and maps a dense entry index
// to its list slot with a binary search.
// Appends the workgroup's surviving tiles (`live` != 0) to `out` (the next
// free slot of its segment) in thread order; returns their number.  Every
// thread of the workgroup must call; s_wave has kThreads/64 entries.
__device__ __forceinline__ unsigned append_local(uint64_t live, uint32_t tile, uint32_t volume,
                                                 ListEntry* __restrict__ out, unsigned* s_wave) {
    const unsigned n = live != 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    unsigned wave_base = 0, total = 0;
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i) {
        wave_base += i < wave ? s_wave[i] : 0u;
        total += s_wave[i];
    }
    if (n) out[wave_base + incl - n] = ListEntry{tile, volume, live};
    __syncthreads();  // s_wave is reused by the next call
    return total;
}

// s_pre[s] = counts[0] + ... + counts[s-1] for s = 0..nseg (nseg <=
// kMaxSegments); returns s_pre[nseg].  Every thread of the workgroup must
// call; s_wave has kThreads/64 entries.
__device__ __forceinline__ unsigned segment_offsets(const unsigned* __restrict__ counts, int nseg,
                                                    unsigned* s_pre, unsigned* s_wave) {
    for (int i = threadIdx.x; i < nseg; i += kThreads) s_pre[i] = counts[i];
    __syncthreads();
    const int per = (nseg + kThreads - 1) / kThreads;
    const int b = threadIdx.x * per, e = min(b + per, nseg);
    unsigned sum = 0;
    for (int i = b; i < e; ++i) sum += s_pre[i];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    unsigned run = incl - sum, total = 0;
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i) {
        run += i < wave ? s_wave[i] : 0u;
        total += s_wave[i];
    }
    for (int i = b; i < e; ++i) {
        const unsigned c = s_pre[i];
        s_pre[i] = run;
        run += c;
    }
    if (threadIdx.x == 0) s_pre[nseg] = total;
    __syncthreads();
    return total;
}

// List slot of dense entry i (< s_pre[nseg]): segment s with s_pre[s] <= i <
// s_pre[s+1] (the last s with s_pre[s] <= i, which skips empty segments),
// then i - s_pre[s] entries into it.  Every lane takes the same number of
// steps.
__device__ __forceinline__ unsigned segment_slot(const unsigned* s_pre, int nseg, unsigned seg_len,
                                                 unsigned i) {
    unsigned lo = 0, len = (unsigned)nseg;
    while (len > 1) {
        const unsigned half = len >> 1;
        if (s_pre[lo + half] <= i) lo += half;
        len -= half;
    }
    return lo * seg_len + (i - s_pre[lo]);
}

// ---- the per-voxel update -------------------------------------------------------
// Projection is branch-free so the 4*NR depth gathers of a frame issue back
// to back; D and W of a row are read only if one of its voxels passes, and
// written once after all frames (frames are applied in order, so the result
// equals one sweep per frame).
//
// COLOR (single-frame calls with a colour image): a voxel updated inside the
// truncation band (-trunc <= sdf <= trunc) also averages the rgb pixel of its
// depth sample into its colour, c' = ((c * wc) + rgb) / (wc + 1) per channel,
// wc' = min(wc + 1, max_weight) -- the running average of the octree's
// add_observation (octree.cu:118-137) with the dense update's capped weight.
// Free-space updates (sdf > trunc) leave the colour alone.
//
// VARW (WEIGHT_BY_VARIANCE): an observation weighs
// variance_weight(d) = sigma(1 m)^2 / sigma(d)^2 instead of 1, so
// d' = ((d * w) + (wo * tsdf)) / (w + wo), w' = min(w + wo, max_weight).
template <class S, class Frames, bool VEC, bool COUNT, bool COLOR = false, bool VARW = false>
__device__ __forceinline__ void integrate_tile(const VolumeParams& vp, const int nframes,
                                               const Frames& frames, const uint64_t live,
                                               const int x0, const int y0, const int z0,
                                               unsigned& cnt, unsigned& touched,
                                               const float* __restrict__ rtab = nullptr) {
    constexpr int NR = S::kRows, NV = 4 * S::kRows;
    float* __restrict__ D = vp.D;
    float* __restrict__ W = vp.W;
    const int lane = threadIdx.x & 63;
    const int x = x0 + 4 * (lane % S::kLanesX);
    const int y = y0 + lane / S::kLanesX;
    if (x >= vp.gx || y >= vp.gy) return;
    const int nvox = VEC ? 4 : min(4, vp.gx - x);
    int base[NR];
    bool row_ok[NR];
    float czw[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int z = z0 + r;
        row_ok[r] = z < vp.gz;
        base[r] = x + vp.gx * (y + vp.gy * min(z, vp.gz - 1));
        czw[r] = voxel_centre(z, vp.vs[2], vp.off[2]);
    }
    const float cyw = voxel_centre(y, vp.vs[1], vp.off[1]);
    float cxw[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) cxw[j] = voxel_centre(x + j, vp.vs[0], vp.off[0]);

    float d[NV], w[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) d[j] = w[j] = 0.0f;
    bool loaded[NR], dirty[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) loaded[r] = dirty[r] = false;
    unsigned upd = 0;  // COUNT: bit j set once voxel j is updated

    // single-frame blocks: a compile-time trip count lets the frame loop
    // unroll (91 -> 59 VGPRs for the plain kernel)
    const int nf = Frames::kMax == 1 ? 1 : nframes;
    for (int f = 0; f < nf; ++f) {
        if (!((live >> f) & 1ull)) continue;
        const FrameParams& F = frames.f[f];
        const float rxy = F.R[1] * cyw, ryy = F.R[4] * cyw, rzy = F.R[7] * cyw;
        bool ok[NV];
        float pz[NV], dep[NV];
        int addr[NV];
        int uu[COLOR ? NV : 1], vv[COLOR ? NV : 1];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int r = j >> 2, i = j & 3;
            const float px = ((F.R[0] * cxw[i] + rxy) + F.R[2] * czw[r]) + F.t[0];
            const float py = ((F.R[3] * cxw[i] + ryy) + F.R[5] * czw[r]) + F.t[1];
            pz[j] = ((F.R[6] * cxw[i] + rzy) + F.R[8] * czw[r]) + F.t[2];
            int u, v;
            const bool in_z = pz[j] > 0.0f;
            const float rz = 1.0f / pz[j];  // one correctly rounded division for both axes
            const bool in_u = pixel_index_rcp(px, vp.fx, pz[j], rz, vp.cx, vp.fcols, u);
            const bool in_v = pixel_index_rcp(py, vp.fy, pz[j], rz, vp.cy, vp.frows, v);
            ok[j] = in_z && in_u && in_v && row_ok[r] && (VEC || i < nvox);
            addr[j] = ok[j] ? v * F.pitch + u : 0;
            if (COLOR) {
                uu[j] = u;
                vv[j] = v;
            }
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) dep[j] = F.depth[addr[j]];
        bool pass[NV], any[NR], band[COLOR ? NV : 1];
        float tsdf[NV];
#pragma unroll
        for (int r = 0; r < NR; ++r) any[r] = false;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const float sdf = dep[j] - pz[j];
            pass[j] = ok[j] && !(dep[j] <= 0.0f || dep[j] > vp.maxd) && (sdf >= -vp.trunc);
            tsdf[j] = sdf > 0.0f ? fminf(sdf, vp.trunc) : sdf;
            any[j >> 2] = any[j >> 2] || pass[j];
            if (COLOR) band[j] = sdf <= vp.trunc;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (any[r] && !loaded[r]) {
                if (VEC) {
                    const float4 d4 = *reinterpret_cast<const float4*>(D + base[r]);
                    const float4 w4 = *reinterpret_cast<const float4*>(W + base[r]);
                    d[4 * r + 0] = d4.x; d[4 * r + 1] = d4.y; d[4 * r + 2] = d4.z; d[4 * r + 3] = d4.w;
                    w[4 * r + 0] = w4.x; w[4 * r + 1] = w4.y; w[4 * r + 2] = w4.z; w[4 * r + 3] = w4.w;
                } else {
                    for (int i = 0; i < nvox; ++i) {
                        d[4 * r + i] = D[base[r] + i];
                        w[4 * r + i] = W[base[r] + i];
                    }
                }
                loaded[r] = true;
            }
        }
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            if (!pass[j]) continue;
            // tsdf_volume.cu:178-186 with the weight cap
            if (VARW) {
                const float wo = variance_weight(vp.noise, vp.wref, dep[j]);
                const float nw = w[j] + wo;
                d[j] = ((d[j] * w[j]) + (wo * tsdf[j])) / nw;
                w[j] = nw < vp.maxw ? nw : vp.maxw;
            } else {
                const float nw = w[j] + 1.0f;
                const float num = (d[j] * w[j]) + tsdf[j];
                // weights are whole counts below the cap: RN(1/nw) from rtab
                // (rtab[i] = 1.0f / (i + 1)) and div_by_rcp, else the division
                const int wi = (int)w[j];
                if (rtab && (float)wi == w[j] && wi < kRcpTable)
                    d[j] = div_by_rcp(num, nw, rtab[wi]);
                else
                    d[j] = num / nw;
                w[j] = nw < vp.maxw ? nw : vp.maxw;
            }
            dirty[j >> 2] = true;
            if (COUNT) {
                ++cnt;
                upd |= 1u << j;
            }
            if (COLOR && band[j]) {
                float4* cp = reinterpret_cast<float4*>(vp.C) + base[j >> 2] + (j & 3);
                float4 c = *cp;
                const uint8_t* px = F.rgb + (int64_t)vv[j] * F.rgb_pitch + 3 * uu[j];
                const float nwc = c.w + 1.0f;
                c.x = ((c.x * c.w) + (float)px[0]) / nwc;
                c.y = ((c.y * c.w) + (float)px[1]) / nwc;
                c.z = ((c.z * c.w) + (float)px[2]) / nwc;
                c.w = nwc < vp.maxw ? nwc : vp.maxw;
                *cp = c;
            }
        }
    }
    if (COUNT) touched += (unsigned)__builtin_popcount(upd);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (!dirty[r]) continue;
        if (VEC) {
            *reinterpret_cast<float4*>(D + base[r]) =
                make_float4(d[4 * r + 0], d[4 * r + 1], d[4 * r + 2], d[4 * r + 3]);
            *reinterpret_cast<float4*>(W + base[r]) =
                make_float4(w[4 * r + 0], w[4 * r + 1], w[4 * r + 2], w[4 * r + 3]);
        } else {
            for (int i = 0; i < nvox; ++i) {
                D[base[r] + i] = d[4 * r + i];
                W[base[r] + i] = w[4 * r + i];
            }
        }
    }
}

// D <- value_d, W <- 0 (tsdf_volume.cu:21-26,230-238), four voxels per thread.
__device__ __forceinline__ void fill4(float* __restrict__ D, float* __restrict__ W, long long i,
                                      long long n, float value_d) {
    if (i + 3 < n && (i & 3) == 0) {
        *reinterpret_cast<float4*>(D + i) = make_float4(value_d, value_d, value_d, value_d);
        *reinterpret_cast<float4*>(W + i) = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        for (long long k = i; k < n && k < i + 4; ++k) {
            D[k] = value_d;
            W[k] = 0.0f;
        }
    }
}

}  // namespace recon
