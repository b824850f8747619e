// tsdf_wrapper_test.cpp -- drives the C++ drop-in (include/recon/tsdf_volume.hpp)
// the way the reference's TSDFSurface::reconstruct_surface does
// (tsdf_surface_reconstruction.cpp:122-153): create a volume, integrate
// frames, extract the mesh, back-project a frame, write the PLY; every result
// is compared bit-for-bit with the CPU oracle (test infrastructure).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "oracle.h"
#include "recon/tsdf_volume.hpp"
#include "scene.h"

static int g_fail = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);             \
            std::fprintf(stderr, "\n");                    \
            ++g_fail;                                      \
        }                                                  \
    } while (0)

// world->camera pose of a camera at c looking at the origin, z up
static std::array<double, 7> look_at(double cx, double cy, double cz) {
    double z[3] = {-cx, -cy, -cz}, n = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (double& v : z) v /= n;
    const double up[3] = {0, 0, 1};
    double x[3] = {z[1] * up[2] - z[2] * up[1], z[2] * up[0] - z[0] * up[2], z[0] * up[1] - z[1] * up[0]};
    n = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (double& v : x) v /= n;
    const double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2], z[0] * x[1] - z[1] * x[0]};
    const double R[9] = {x[0], x[1], x[2], y[0], y[1], y[2], z[0], z[1], z[2]};
    // quaternion (w >= 0) of R
    double w = std::sqrt(std::fmax(0.0, 1.0 + R[0] + R[4] + R[8])) / 2;
    double qx = std::sqrt(std::fmax(0.0, 1.0 + R[0] - R[4] - R[8])) / 2;
    double qy = std::sqrt(std::fmax(0.0, 1.0 - R[0] + R[4] - R[8])) / 2;
    double qz = std::sqrt(std::fmax(0.0, 1.0 - R[0] - R[4] + R[8])) / 2;
    qx = std::copysign(qx, R[7] - R[5]);
    qy = std::copysign(qy, R[2] - R[6]);
    qz = std::copysign(qz, R[3] - R[1]);
    const double c[3] = {cx, cy, cz};
    std::array<double, 7> p{w, qx, qy, qz, 0, 0, 0};
    for (int i = 0; i < 3; ++i) p[4 + i] = -(R[3 * i] * c[0] + R[3 * i + 1] * c[1] + R[3 * i + 2] * c[2]);
    return p;
}

int main() {
    const int cols = 96, rows = 72, nframes = 4;
    const double fx = 80.0, fy = 80.0, ccx = 47.5, ccy = 35.5;
    recon_scene_prim prims[2] = {};
    prims[0].kind = RECON_PRIM_SPHERE;
    prims[0].p[3] = 0.25;
    prims[1].kind = RECON_PRIM_PLANE_Z;
    prims[1].p[0] = -0.3;
    std::vector<std::array<double, 7>> poses;
    std::vector<double> flat;
    for (int f = 0; f < nframes; ++f) {
        const double th = 0.3 * f;
        poses.push_back(look_at(std::cos(th), std::sin(th), 0.3));
        flat.insert(flat.end(), poses.back().begin(), poses.back().end());
    }
    recon_scene_noise noise{0.001, 0.0, 0.02, 99};
    std::vector<float> depth((size_t)nframes * rows * cols);
    CHECK(recon_scene_render(prims, 2, flat.data(), nullptr, nframes, fx, fy, ccx, ccy, cols, rows,
                             &noise, depth.data(), 0) == 0, "render");
    float* ddepth = nullptr;
    CHECK(hipMalloc((void**)&ddepth, depth.size() * 4) == hipSuccess, "hipMalloc");
    CHECK(hipMemcpy(ddepth, depth.data(), depth.size() * 4, hipMemcpyHostToDevice) == hipSuccess, "upload");

    cuphoto::TSDFVolumeConfig cfg{};
    cfg.voxel_grid_size = dim3(60, 52, 44);
    cfg.physical_size = make_float3(0.9f, 0.78f, 0.66f);
    cfg.global_offset = make_float3(-0.45f, -0.39f, -0.33f);
    cfg.max_weight = 3.0f;
    cfg.focal_x = (float)fx;
    cfg.focal_y = (float)fy;
    cfg.cx = (float)ccx;
    cfg.cy = (float)ccy;
    cfg.camera_width = cols;
    cfg.camera_height = rows;
    cfg.truncation = 0.04f;
    cuphoto::TSDFVolume vol(cfg);
    CHECK(vol.ok(), "create");
    cuphoto::PtrStepSzb no_color;
    for (int f = 0; f < nframes; ++f) {
        cuphoto::PtrStepSzf d;
        d.data = ddepth + (size_t)f * rows * cols;
        d.step = cols * 4;
        d.rows = rows;
        d.cols = cols;
        CHECK(vol.integrate(no_color, d, poses[f]), "integrate frame %d", f);
    }
    // error behaviour: a null depth map fails with false (tsdf_volume.cu:260-263)
    cuphoto::PtrStepSzf null_depth;
    null_depth.rows = rows;
    null_depth.cols = cols;
    null_depth.step = cols * 4;
    CHECK(!vol.integrate(no_color, null_depth, poses[0]), "null depth must fail");

    const dim3 g = vol.voxel_grid_size();
    const size_t nvox = (size_t)g.x * g.y * g.z;
    std::vector<float> D(nvox), W(nvox);
    CHECK(hipMemcpy(D.data(), vol.voxel_distances_data(), nvox * 4, hipMemcpyDeviceToHost) == hipSuccess, "D");
    CHECK(hipMemcpy(W.data(), vol.voxel_weights_data(), nvox * 4, hipMemcpyDeviceToHost) == hipSuccess, "W");

    oracle_volume ov{};
    ov.grid[0] = g.x; ov.grid[1] = g.y; ov.grid[2] = g.z;
    const float3 vs = vol.scale_voxel_size();
    ov.voxel_size[0] = vs.x; ov.voxel_size[1] = vs.y; ov.voxel_size[2] = vs.z;
    ov.offset[0] = cfg.global_offset.x; ov.offset[1] = cfg.global_offset.y; ov.offset[2] = cfg.global_offset.z;
    ov.truncation = vol.truncation();
    ov.max_weight = cfg.max_weight;
    ov.max_depth = 24.0f;
    ov.fx = cfg.focal_x; ov.fy = cfg.focal_y; ov.cx = cfg.cx; ov.cy = cfg.cy;
    std::vector<float> Do(nvox), Wo(nvox);
    oracle_tsdf_init(&ov, Do.data(), Wo.data());
    int64_t updates = 0;
    for (int f = 0; f < nframes; ++f)
        updates += oracle_tsdf_integrate_f32(&ov, Do.data(), Wo.data(), depth.data() + (size_t)f * rows * cols,
                                             rows, cols, cols, poses[f].data(), nullptr, 0);
    CHECK(updates > 1000, "updates %lld", (long long)updates);
    CHECK(std::memcmp(D.data(), Do.data(), nvox * 4) == 0, "TSDF differs from the oracle");
    CHECK(std::memcmp(W.data(), Wo.data(), nvox * 4) == 0, "weights differ from the oracle");

    std::vector<float3> verts;
    std::vector<int3> tris;
    verts.push_back(make_float3(7.f, 7.f, 7.f));  // generate_triangular_surface appends
    cuphoto::generate_triangular_surface(vol, verts, tris);
    const int64_t nt = oracle_mc_count(Do.data(), ov.grid, nullptr, 0);
    CHECK((int64_t)tris.size() == nt && nt > 100, "triangles %zu vs %lld", tris.size(), (long long)nt);
    std::vector<float> ovs(9 * (size_t)nt);
    std::vector<int32_t> ots(3 * (size_t)nt);
    oracle_mc_extract(Do.data(), ov.grid, ov.voxel_size, ov.offset, ovs.data(), ots.data(), nullptr,
                      nullptr, nt, 0);
    CHECK(verts.size() == 1 + 3 * (size_t)nt, "vertex count");
    CHECK(std::memcmp(verts.data() + 1, ovs.data(), ovs.size() * 4) == 0, "vertices differ from the oracle");
    bool idx_ok = true;
    for (size_t t = 0; t < tris.size(); ++t)
        idx_ok &= tris[t].x == (int)(1 + 3 * t) && tris[t].y == (int)(2 + 3 * t) && tris[t].z == (int)(3 + 3 * t);
    CHECK(idx_ok, "triangle indices");

    const std::string ply = "/tmp/recon_wrapper_test_mesh.ply";
    CHECK(cuphoto::write_mesh_to_ply(ply, verts, tris), "ply");
    std::ifstream in(ply);
    std::string line;
    std::getline(in, line);
    CHECK(line == "ply", "ply header");

    // depth -> points for frame 2 into slot 1 of a two-frame cloud
    const std::array<double, 9> K = {fx, 0, 0, 0, fy, 0, ccx, ccy, 1};  // column-major
    cuphoto::PointCloud pc(K, 2 * (size_t)rows * cols);
    cuphoto::PtrStepSzf d2;
    d2.data = ddepth + 2 * (size_t)rows * cols;
    d2.step = cols * 4;
    d2.rows = rows;
    d2.cols = cols;
    CHECK(pc.extract_points(d2, no_color, poses[2], 1), "extract_points");
    std::vector<float> xyz(3 * (size_t)rows * cols), oxyz(3 * (size_t)rows * cols);
    std::vector<uint8_t> valid((size_t)rows * cols), ovalid((size_t)rows * cols);
    CHECK(hipMemcpy(xyz.data(), pc.xyz() + 3 * (size_t)rows * cols, xyz.size() * 4, hipMemcpyDeviceToHost) == hipSuccess, "xyz");
    CHECK(hipMemcpy(valid.data(), pc.valid() + (size_t)rows * cols, valid.size(), hipMemcpyDeviceToHost) == hipSuccess, "valid");
    oracle_depth_to_points(depth.data() + 2 * (size_t)rows * cols, rows, cols, cols, fx, fy, ccx, ccy,
                           poses[2].data(), nullptr, 0, oxyz.data(), ovalid.data(), nullptr);
    CHECK(std::memcmp(xyz.data(), oxyz.data(), xyz.size() * 4) == 0, "points differ from the oracle");
    CHECK(valid == ovalid, "validity differs");
    CHECK(!pc.extract_points(d2, no_color, poses[2], 2), "frame 2 does not fit and must fail");

    (void)hipFree(ddepth);
    if (g_fail) {
        std::fprintf(stderr, "%d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("wrapper ok: %lld voxel updates, %zu triangles, bit-exact vs oracle\n",
                (long long)updates, tris.size());
    return 0;
}
