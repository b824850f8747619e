// color.hip -- colour of the fused surface at given points for gfx950.
//
// No counterpart on the reference's dense path, whose integrate accepts a
// colour image and ignores it (tsdf_volume.cu:257-290); its octree keeps a
// per-node colour average (octree.cu:118-137).  Here a volume with colour
// enabled holds (r, g, b, wc) per voxel (the colour integrate is in
// integrate_device.hpp), and recon_tsdf_sample_colors reads it at mesh
// vertices or any other points: the trilinear support of the TSDF sampler
// (tri_support: tsdf_volume.cu:46-104 with the weights of neighbours 6 and 7
// in their own places), each neighbour weighted by its trilinear coefficient
// times its colour weight,
//     c = sum(coef_i * wc_i * c_i) / sum(coef_i * wc_i),
// summed in neighbour order 0..7, rounded to the nearest integer in 0..255;
// (0, 0, 0) where no neighbour has colour or the point has no support.
#pragma clang fp contract(off)

#include "diff_device.hpp"

#include <cmath>

namespace recon {

__global__ __launch_bounds__(256) void sample_colors_kernel(const VolumeParams vp,
                                                            const float4* __restrict__ C,
                                                            const float* __restrict__ points,
                                                            int64_t n, uint8_t* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const float p[3] = {points[3 * k], points[3 * k + 1], points[3 * k + 2]};
    int64_t ids[8];
    float coef[8];
    float r = 0.0f, g = 0.0f, b = 0.0f, wsum = 0.0f;
    if (tri_support(vp, p, ids, coef)) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float4 c = C[ids[i]];
            const float a = coef[i] * c.w;
            r = r + a * c.x;
            g = g + a * c.y;
            b = b + a * c.z;
            wsum = wsum + a;
        }
    }
    uint8_t q[3] = {0, 0, 0};
    if (wsum > 0.0f) {
        const float v[3] = {r / wsum, g / wsum, b / wsum};
#pragma unroll
        for (int i = 0; i < 3; ++i) q[i] = (uint8_t)fminf(fmaxf(rintf(v[i]), 0.0f), 255.0f);
    }
    out[3 * k] = q[0];
    out[3 * k + 1] = q[1];
    out[3 * k + 2] = q[2];
}

}  // namespace recon

using namespace recon;

extern "C" int recon_tsdf_sample_colors(recon_tsdf* h, const float* points, int64_t num_points,
                                        uint8_t* colors_out, recon_stream_t stream) {
    clear_error();
    const char* fn = "recon_tsdf_sample_colors";
    if (!h) return fail(RECON_ERR_NULL, "%s: handle is NULL", fn);
    if (!h->C) return fail(RECON_ERR_CONFIG, "%s: colour is not enabled on this volume", fn);
    if (num_points < 0) return fail(RECON_ERR_SIZE, "%s: num_points < 0", fn);
    if (num_points == 0) return RECON_OK;
    if (!points || !colors_out) return fail(RECON_ERR_NULL, "%s: points or output is NULL", fn);
    if (num_points >= (int64_t(1) << 31)) return fail(RECON_ERR_RANGE, "%s: too many points", fn);
    const int rc = require_gfx950();
    if (rc) return rc;
    hipLaunchKernelGGL(sample_colors_kernel, dim3((unsigned)((num_points + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, h->vp, reinterpret_cast<const float4*>(h->C), points,
                       num_points, colors_out);
    RECON_HIP_TRY(hipGetLastError(), "sample_colors_kernel launch");
    return RECON_OK;
}
