// diff_device.hpp -- device code of the differentiable integrate (an
// extension of cu_tsdf_integrate, cu-Reconstruction
// modules/src/cuda/tsdf_volume.cu:144-190, with bilinear depth sampling), of
// the surface loss on the trilinear sampler (interpolate_trilinearly,
// tsdf_volume.cu:46-104) and of the adjoint of both with respect to per-frame
// se(3) poses (left perturbation T <- Exp(xi) T, xi = (rho, omega), so a
// camera point moves by rho + omega x p) and to the depth maps.
//
// Bilinear sampling: depth[j][i] sits at pixel coordinates (i, j) (the ray of
// camera_util.cu:21-31).  A voxel samples frame f when its camera z > 0, its
// continuous pixel coordinates s_u = ((x*fx)/z) + cx, s_v = ((y*fy)/z) + cy
// (camera_util.cu:7-19 before the truncation) satisfy 0 <= s < size - 1 (the
// 2x2 footprint lies in the image), and all four taps are valid depths
// (0 < d <= max_depth).  Otherwise the update is the plain one: skip when
// sdf < -trunc, tsdf = min(sdf, trunc) for sdf > 0, weight cap.
//
// The forward kernel and the adjoint evaluate every decision with the same
// float32 expressions (bi_project, bi_depth), so the adjoint replays exactly
// the updates the forward made.
#pragma once

#include "integrate_device.hpp"

namespace recon {

struct BiProj {
    float px, py, pz;  // camera point
    float a, b;        // fractional pixel offsets of the sample
    int addr;          // top-left tap j0 * pitch + i0 (0 when !in)
    bool in;           // z > 0 and the 2x2 footprint inside the image
};

__device__ __forceinline__ BiProj bi_project(const VolumeParams& vp, const FrameParams& F, float cx,
                                             float cy, float cz) {
    BiProj s;
    s.px = ((F.R[0] * cx + F.R[1] * cy) + F.R[2] * cz) + F.t[0];
    s.py = ((F.R[3] * cx + F.R[4] * cy) + F.R[5] * cz) + F.t[1];
    s.pz = ((F.R[6] * cx + F.R[7] * cy) + F.R[8] * cz) + F.t[2];
    // one correctly rounded 1/z serves both quotients (div_by_rcp equals the
    // division for normal-range quotients; outside that range both differ
    // from cx, cy by less than half their ulp or are invalid alike)
    const float rz = 1.0f / s.pz;
    const float su = div_by_rcp(s.px * vp.fx, s.pz, rz) + vp.cx;
    const float sv = div_by_rcp(s.py * vp.fy, s.pz, rz) + vp.cy;
    s.in = s.pz > 0.0f && su >= 0.0f && su < vp.fcols - 1.0f && sv >= 0.0f && sv < vp.frows - 1.0f;
    const int i0 = s.in ? (int)su : 0, j0 = s.in ? (int)sv : 0;
    s.a = su - (float)i0;
    s.b = sv - (float)j0;
    s.addr = j0 * F.pitch + i0;
    return s;
}

__device__ __forceinline__ bool tap_valid(float d, float maxd) { return d > 0.0f && d <= maxd; }

// Bilinear interpolation of the four taps in a fixed order.
__device__ __forceinline__ float bi_depth(float a, float b, float d00, float d10, float d01,
                                          float d11) {
    const float ia = 1.0f - a, ib = 1.0f - b;
    return (((ia * ib) * d00 + (a * ib) * d10) + (ia * b) * d01) + (a * b) * d11;
}

// ---- forward: one wave-tile, frames in order ---------------------------------
// The structure of integrate_tile: lane l owns voxels x..x+3 of row y for
// S::kRows z; the 4 * 4 * kRows depth taps of a frame issue back to back;
// D and W are read when a row first passes and written once.
template <class S, class Frames, bool VEC>
__device__ __forceinline__ void diff_integrate_tile(const VolumeParams& vp, const int nframes,
                                                    const Frames& frames, const uint64_t live,
                                                    const int x0, const int y0, const int z0,
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

    for (int f = 0; f < nframes; ++f) {
        if (!((live >> f) & 1ull)) continue;
        const FrameParams& F = frames.f[f];
        BiProj pr[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int r = j >> 2, i = j & 3;
            pr[j] = bi_project(vp, F, cxw[i], cyw, czw[r]);
            pr[j].in = pr[j].in && row_ok[r] && (VEC || i < nvox);
            pr[j].addr = pr[j].in ? pr[j].addr : 0;
        }
        float t00[NV], t10[NV], t01[NV], t11[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            // the depth maps are device memory: global loads, not flat ones
            const __attribute__((address_space(1))) float* q =
                (const __attribute__((address_space(1))) float*)(F.depth + pr[j].addr);
            t00[j] = q[0];
            t10[j] = q[1];
            t01[j] = q[F.pitch];
            t11[j] = q[F.pitch + 1];
        }
        bool pass[NV], any[NR];
        float tsdf[NV];
#pragma unroll
        for (int r = 0; r < NR; ++r) any[r] = false;
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const bool valid = pr[j].in && tap_valid(t00[j], vp.maxd) && tap_valid(t10[j], vp.maxd) &&
                               tap_valid(t01[j], vp.maxd) && tap_valid(t11[j], vp.maxd);
            const float sdf = bi_depth(pr[j].a, pr[j].b, t00[j], t10[j], t01[j], t11[j]) - pr[j].pz;
            pass[j] = valid && (sdf >= -vp.trunc);
            tsdf[j] = sdf > 0.0f ? fminf(sdf, vp.trunc) : sdf;
            any[j >> 2] = any[j >> 2] || pass[j];
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
            const float nw = w[j] + 1.0f;
            const float num = (d[j] * w[j]) + tsdf[j];
            // whole weights below kRcpTable: RN(1/nw) from rtab and div_by_rcp
            const int wi = (int)w[j];
            if (rtab && (float)wi == w[j] && wi < kRcpTable)
                d[j] = div_by_rcp(num, nw, rtab[wi]);
            else
                d[j] = num / nw;
            w[j] = nw < vp.maxw ? nw : vp.maxw;
            dirty[j >> 2] = true;
        }
    }
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

// ---- the trilinear sampler ------------------------------------------------------
// interpolate_trilinearly (tsdf_volume.cu:46-104) on the grid-local point
// q = p - offset: a coordinate more than BOUNDARY_EPS below the grid moves
// onto its face, one less than BOUNDARY_EPS above it moves inside
// (:53-64); voxel = floor(q / vs) (:28-35), lower corner one below when q is
// under the voxel's centre, clamped at 0 (:73-78); weights from
// (q - centre(lower)) / vs (:79-81,94-101) in the neighbour order of
// :85-92.  Two departures: neighbour 6 = (1,1,1) gets x*y*z and neighbour
// 7 = (0,1,1) gets (1-x)*y*z (the reference swaps the two weights), and a
// support that would leave the grid is reported (false) instead of read.
__device__ __forceinline__ bool tri_support(const VolumeParams& vp, const float p[3],
                                            int64_t ids[8], float coef[8]) {
    const int g[3] = {vp.gx, vp.gy, vp.gz};
    int lo[3];
    float t[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float q = p[k] - vp.off[k];
        const float bound = (float)g[k] * vp.vs[k];
        if (q < -1e-4f) q = 0.0f;
        if ((q > bound) && (q - bound < 1e-4f)) q = bound - 1e-4f;
        // beyond the grid (or NaN): no support; this also keeps the index
        // conversion below in range
        if (!(q <= bound)) return false;
        const int vox = (int)floorf(q / vp.vs[k]);
        const float ct = ((float)vox + 0.5f) * vp.vs[k];
        int l = q < ct ? vox - 1 : vox;
        l = max(l, 0);
        if (!(l + 1 <= g[k] - 1)) return false;
        lo[k] = l;
        const float lct = ((float)l + 0.5f) * vp.vs[k];
        t[k] = (q - lct) / vp.vs[k];
    }
    const int64_t dx = 1, dy = vp.gx, dz = (int64_t)vp.gx * vp.gy;
    ids[0] = lo[0] + (int64_t)vp.gx * (lo[1] + (int64_t)vp.gy * lo[2]);
    ids[1] = ids[0] + dx;
    ids[2] = ids[1] + dz;
    ids[3] = ids[0] + dz;
    ids[4] = ids[0] + dy;
    ids[5] = ids[1] + dy;
    ids[6] = ids[2] + dy;
    ids[7] = ids[3] + dy;
    const float x = t[0], y = t[1], z = t[2];
    coef[0] = (1 - x) * (1 - y) * (1 - z);
    coef[1] = x * (1 - y) * (1 - z);
    coef[2] = x * (1 - y) * z;
    coef[3] = (1 - x) * (1 - y) * z;
    coef[4] = (1 - x) * y * (1 - z);
    coef[5] = x * y * (1 - z);
    coef[6] = x * y * z;
    coef[7] = (1 - x) * y * z;
    return true;
}

// ---- reductions ------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// ---- adjoint, lanes as frames ----------------------------------------------------
// For up to 64 frames a wave holds S = 64 / F segments of F lanes; lane f of
// segment s evaluates frame f of the segment's voxel once.  The running
// average of a voxel is D' = a D + b tsdf, a = Wb/(Wb+1), b = 1/(Wb+1),
// Wb = min(updates before the frame, cap): the updates before frame f are
// the set bits of the segment's ballot below the lane, and dL/dD after frame
// f is lambda times the product of a over the segment's later frames (a
// suffix-product scan).  With g = that * b (0 when the frame clamps at
// +trunc) and sdf = d(s_u, s_v) - z:
//   dsdf/dp = (dd/du fx/z, dd/dv fy/z, -(dd/du fx x + dd/dv fy y)/z^2 - 1),
//   dL/drho += g dsdf/dp,  dL/domega += p x (g dsdf/dp),
//   dL/d(tap) += g * (bilinear weight of the tap).
struct LaneLayout {
    int F, S, seg, f;
    bool ok;             // the lane belongs to a segment
    uint64_t seg_mask;   // lanes of this lane's segment
    uint64_t below;      // lanes below this one
};

__device__ __forceinline__ LaneLayout lane_layout(int F) {
    LaneLayout ly;
    const int lane = threadIdx.x & 63;
    ly.F = F;
    ly.S = 64 / F;
    ly.seg = lane / F;
    ly.f = lane - ly.seg * F;
    ly.ok = ly.seg < ly.S;
    ly.seg_mask = (F == 64 ? ~0ull : ((1ull << F) - 1ull)) << (ly.ok ? ly.seg * F : 0);
    ly.below = (1ull << lane) - 1ull;
    return ly;
}

// Depth gradients are summed exactly: each tap's contribution c is added to
// an int64 accumulator as llrint(c * scale), scale = 2^(41 - e) for the
// largest |dL/dD| in [2^e, 2^(e+1)) -- a contribution is at most that
// value (the running-average factors and bilinear weights are <= 1), so
// 2^20 contributions to one pixel cannot overflow, and the integer sum does
// not depend on the order of the atomics.
__device__ __forceinline__ double depth_grad_scale(unsigned maxbits) {
    const float m = __uint_as_float(maxbits);
    return m > 0.0f ? ldexp(1.0, 41 - ilogbf(m)) : 1.0;
}

__device__ __forceinline__ void depth_grad_add(unsigned long long* o, float c, double scale) {
    const long long q = llrint((double)c * scale);
    if (q) atomicAdd(o, (unsigned long long)q);
}

// One pass of the wave: this lane's segment holds voxel v (local index, or
// -1) with dL/dD = lam; L is the lane's frame.  gd_frame (may be null): the
// lane frame's depth-gradient accumulators, one per depth element, scaled
// by gd_scale.  Every lane must call.
__device__ __forceinline__ void adjoint_lane_pass(const VolumeParams& vp, const LaneLayout& ly,
                                                  const FrameParams& L, int64_t v, float lam,
                                                  unsigned long long* gd_frame, double gd_scale,
                                                  double acc[6]) {
    const int64_t N = (int64_t)vp.gx * vp.gy * vp.gz;
    const bool act = ly.ok && v >= 0 && v < N && lam != 0.0f && isfinite(lam);
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (act) {
        const int64_t x = v % vp.gx, yz = v / vp.gx;
        cx = voxel_centre((int)x, vp.vs[0], vp.off[0]);
        cy = voxel_centre((int)(yz % vp.gy), vp.vs[1], vp.off[1]);
        cz = voxel_centre((int)(yz / vp.gy), vp.vs[2], vp.off[2]);
    }
    bool pass = false;
    BiProj s;
    s.px = s.py = s.pz = s.a = s.b = 0.0f;
    s.addr = 0;
    float t00 = 0.0f, t10 = 0.0f, t01 = 0.0f, t11 = 0.0f, sdf = 0.0f;
    if (act) {
        s = bi_project(vp, L, cx, cy, cz);
        if (s.in) {
            const float* q = L.depth + s.addr;
            t00 = q[0];
            t10 = q[1];
            t01 = q[L.pitch];
            t11 = q[L.pitch + 1];
            const bool valid = tap_valid(t00, vp.maxd) && tap_valid(t10, vp.maxd) &&
                               tap_valid(t01, vp.maxd) && tap_valid(t11, vp.maxd);
            sdf = bi_depth(s.a, s.b, t00, t10, t01, t11) - s.pz;
            pass = valid && sdf >= -vp.trunc;
        }
    }
    const uint64_t ballot = __ballot(pass);
    const int before = __popcll(ballot & ly.seg_mask & ly.below);
    float a = 1.0f, b = 0.0f;
    if (pass) {
        const float wb = fminf((float)before, vp.maxw);
        a = wb / (wb + 1.0f);
        b = 1.0f / (wb + 1.0f);
    }
    // P = product of a over the later frames of the segment
    const float nxt = __shfl_down(a, 1, 64);
    float P = (ly.f + 1 < ly.F) ? nxt : 1.0f;
    for (int d = 1; d < ly.F; d <<= 1) {
        const float y = __shfl_down(P, d, 64);
        if (ly.f + d < ly.F) P = P * y;
    }
    if (pass && !(sdf > 0.0f && sdf >= vp.trunc)) {
        const float g = (lam * P) * b;
        if (g != 0.0f) {
            const float ia = 1.0f - s.a, ib = 1.0f - s.b;
            const float dd_du = ib * (t10 - t00) + s.b * (t11 - t01);
            const float dd_dv = ia * (t01 - t00) + s.a * (t11 - t10);
            const float iz = 1.0f / s.pz;
            const float gx = ((g * dd_du) * vp.fx) * iz;
            const float gy = ((g * dd_dv) * vp.fy) * iz;
            const float gz = -((gx * s.px + gy * s.py) * iz) - g;
            acc[0] += gx;
            acc[1] += gy;
            acc[2] += gz;
            acc[3] += (double)s.py * gz - (double)s.pz * gy;
            acc[4] += (double)s.pz * gx - (double)s.px * gz;
            acc[5] += (double)s.px * gy - (double)s.py * gx;
            if (gd_frame) {
                unsigned long long* o = gd_frame + s.addr;
                depth_grad_add(o, g * (ia * ib), gd_scale);
                depth_grad_add(o + 1, g * (s.a * ib), gd_scale);
                depth_grad_add(o + L.pitch, g * (ia * s.b), gd_scale);
                depth_grad_add(o + L.pitch + 1, g * (s.a * s.b), gd_scale);
            }
        }
    }
}

// Sum of x over the segments of the wave, in segment order, valid in
// segment 0.  Every lane must call.
__device__ __forceinline__ double segment_sum(double x, const LaneLayout& ly) {
    const int lane = threadIdx.x & 63;
    double sum = x;
    for (int k = 1; k < ly.S; ++k) {
        const double o = __shfl(x, lane + k * ly.F, 64);
        if (ly.seg == 0) sum += o;
    }
    return sum;
}

}  // namespace recon
