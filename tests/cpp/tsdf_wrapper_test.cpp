// tsdf_wrapper_test.cpp -- drives the C++ drop-in (include/recon/tsdf_volume.hpp)
// the way the reference's TSDFSurface::reconstruct_surface does
// (tsdf_surface_reconstruction.cpp:122-153): create a volume, integrate
// frames, extract the mesh, back-project frames, render the point cloud,
// write the PLY; every result is compared bit-for-bit with the CPU oracle
// (test infrastructure).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "oracle.h"
#include "recon/tsdf_surface.hpp"
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
    {
        cuphoto::PtrStepSzb small_color{(unsigned char*)ddepth, 3 * (size_t)cols, rows - 1, cols};
        cuphoto::PtrStepSzf d0{ddepth, (size_t)cols * 4, rows, cols};
        CHECK(!vol.integrate(small_color, d0, poses[0]), "a colour map smaller than the depth must fail");
    }

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

    // transformation_voxels_data: the voxel centre and the identity quaternion
    // per voxel (tsdf_volume.cu:106-127); the same pointer on a second call
    {
        using TV = cuphoto::TSDFVolume::TransformationVoxel;
        const TV* tv = vol.transformation_voxels_data();
        CHECK(tv != nullptr && tv == vol.transformation_voxels_data(), "transformation_voxels_data");
        std::vector<TV> T(nvox);
        if (tv)
            CHECK(hipMemcpy(T.data(), tv, nvox * sizeof(TV), hipMemcpyDeviceToHost) == hipSuccess, "T");
        bool tv_ok = tv != nullptr;
        for (size_t i = 0; tv_ok && i < nvox; ++i) {
            float c[3];
            oracle_voxel_centre_f32(&ov, (int32_t)(i % g.x), (int32_t)((i / g.x) % g.y),
                                    (int32_t)(i / ((size_t)g.x * g.y)), c);
            const float want[7] = {c[0], c[1], c[2], 0.f, 0.f, 0.f, 1.f};
            tv_ok = std::memcmp(&T[i], want, sizeof(want)) == 0;
            if (!tv_ok) std::fprintf(stderr, "transformation voxel %zu differs\n", i);
        }
        CHECK(tv_ok, "transformation voxels differ from the oracle's centres");
    }

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
    cuphoto::cudaPointCloud::Ptr pc = cuphoto::cudaPointCloud::create(K, 2 * (size_t)rows * cols);
    CHECK(pc && pc->get_total_num_points() == 2 * (size_t)rows * cols && pc->get_num_points() == 0,
          "cudaPointCloud::create");
    cuphoto::PtrStepSzf d2;
    d2.data = ddepth + 2 * (size_t)rows * cols;
    d2.step = cols * 4;
    d2.rows = rows;
    d2.cols = cols;
    CHECK(pc->extract_points(d2, no_color, poses[2], 1), "extract_points");
    std::vector<float> xyz(3 * (size_t)rows * cols), oxyz(3 * (size_t)rows * cols);
    std::vector<uint8_t> valid((size_t)rows * cols), ovalid((size_t)rows * cols);
    CHECK(hipMemcpy(xyz.data(), pc->xyz() + 3 * (size_t)rows * cols, xyz.size() * 4, hipMemcpyDeviceToHost) == hipSuccess, "xyz");
    CHECK(hipMemcpy(valid.data(), pc->valid() + (size_t)rows * cols, valid.size(), hipMemcpyDeviceToHost) == hipSuccess, "valid");
    oracle_depth_to_points(depth.data() + 2 * (size_t)rows * cols, rows, cols, cols, fx, fy, ccx, ccy,
                           poses[2].data(), nullptr, 0, oxyz.data(), ovalid.data(), nullptr);
    CHECK(std::memcmp(xyz.data(), oxyz.data(), xyz.size() * 4) == 0, "points differ from the oracle");
    CHECK(valid == ovalid, "validity differs");
    CHECK(!pc->extract_points(d2, no_color, poses[2], 2), "frame 2 does not fit and must fail");

    // z-buffer render (project_to_color_depth / project_to_depth) of a
    // two-frame cloud: slot 0 gets frame 1 in world coordinates
    {
        const std::array<double, 7>& P = poses[1];
        const double w = P[0], x = P[1], y = P[2], z = P[3];
        const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                             2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                             2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
        std::array<double, 7> inv{w, -x, -y, -z, 0, 0, 0};
        for (int i = 0; i < 3; ++i)
            inv[4 + i] = -(R[i] * P[4] + R[3 + i] * P[5] + R[6 + i] * P[6]);  // -R^T t
        cuphoto::PtrStepSzf d1{ddepth + (size_t)rows * cols, (size_t)cols * 4, rows, cols};
        CHECK(pc->extract_points(d1, no_color, inv, 0), "extract_points slot 0");
    }
    const size_t np = pc->get_num_points(), n1 = (size_t)rows * cols;
    CHECK(np == 2 * n1, "points %zu", np);
    std::vector<float> all_xyz(3 * np);
    std::vector<uint8_t> all_valid(np), all_rgb(3 * np);
    CHECK(hipMemcpy(all_xyz.data(), pc->xyz(), 12 * np, hipMemcpyDeviceToHost) == hipSuccess, "xyz");
    CHECK(hipMemcpy(all_valid.data(), pc->valid(), np, hipMemcpyDeviceToHost) == hipSuccess, "valid");
    CHECK(hipMemcpy(all_rgb.data(), pc->colors(), 3 * np, hipMemcpyDeviceToHost) == hipSuccess, "rgb");
    const size_t dstep = 4 * ((size_t)cols + 7), cstep = 3 * (size_t)cols + 5;
    float* rdepth = nullptr;
    uint8_t* rcolor = nullptr;
    CHECK(hipMalloc((void**)&rdepth, dstep * rows) == hipSuccess, "hipMalloc");
    CHECK(hipMalloc((void**)&rcolor, cstep * rows) == hipSuccess, "hipMalloc");
    cuphoto::PtrStepSzf rd{rdepth, dstep, rows, cols};
    cuphoto::PtrStepSzb rc{rcolor, cstep, rows, cols};
    CHECK(pc->project_to_color_depth(rc, rd, poses[3]), "project_to_color_depth");
    std::vector<float> gdepth(dstep / 4 * rows), odepth(dstep / 4 * rows);
    std::vector<uint8_t> gcolor(cstep * rows), ocolor(cstep * rows);
    CHECK(hipMemcpy(gdepth.data(), rdepth, dstep * rows, hipMemcpyDeviceToHost) == hipSuccess, "dl");
    CHECK(hipMemcpy(gcolor.data(), rcolor, cstep * rows, hipMemcpyDeviceToHost) == hipSuccess, "dl");
    oracle_render_points(all_xyz.data(), all_valid.data(), all_rgb.data(), (int64_t)np,
                         poses[3].data(), (float)fx, (float)fy, (float)ccx, (float)ccy, rows, cols,
                         (int64_t)(dstep / 4), odepth.data(), ocolor.data(), (int64_t)cstep);
    size_t hits = 0;
    bool same = true;
    for (int r = 0; r < rows; ++r) {
        same &= std::memcmp(&gdepth[r * (dstep / 4)], &odepth[r * (dstep / 4)], 4 * (size_t)cols) == 0;
        same &= std::memcmp(&gcolor[r * cstep], &ocolor[r * cstep], 3 * (size_t)cols) == 0;
        for (int c = 0; c < cols; ++c) hits += odepth[r * (dstep / 4) + c] > 0.f;
    }
    CHECK(same, "render differs from the oracle");
    CHECK(hits > n1 / 4, "render hits %zu", hits);
    CHECK(pc->project_to_depth(rd, poses[3]), "project_to_depth");
    CHECK(hipMemcpy(gdepth.data(), rdepth, dstep * rows, hipMemcpyDeviceToHost) == hipSuccess, "dl");
    same = true;
    for (int r = 0; r < rows; ++r)
        same &= std::memcmp(&gdepth[r * (dstep / 4)], &odepth[r * (dstep / 4)], 4 * (size_t)cols) == 0;
    CHECK(same, "project_to_depth differs from the oracle");
    // get_points(): the reference's packed 33-byte records, readable on the host
    {
        const cuphoto::cudaPointCloud::Vertex* v = pc->get_points();
        CHECK(v != nullptr && pc->get_size() == 33 * np, "get_points / get_size");
        bool vok = v != nullptr;
        for (size_t i = 0; vok && i < np; ++i) {
            vok &= std::memcmp(&v[i].pos, &all_xyz[3 * i], 12) == 0;
            vok &= v[i].color.x == all_rgb[3 * i] && v[i].color.y == all_rgb[3 * i + 1] &&
                   v[i].color.z == all_rgb[3 * i + 2];
            vok &= v[i].normal.x == 0.f && v[i].normal.w == 0.f && v[i].classID == 0;
        }
        CHECK(vok, "packed vertices differ from the planes");
        CHECK(pc->get_vertex(n1 + 5) == v + n1 + 5, "get_vertex");
        const std::array<double, 9> Kc = pc->get_camera_parameters();
        CHECK(Kc == K, "get_camera_parameters");
    }
    cuphoto::PtrStepSzf null_rd{nullptr, dstep, rows, cols};
    CHECK(!pc->project_to_depth(null_rd, poses[3]), "null depth must fail");
    (void)hipFree(rdepth);
    (void)hipFree(rcolor);

    // TSDFSurface::reconstruct_surface(cloud): render at the global pose,
    // integrate, marching cubes (tsdf_surface_reconstruction.cpp:94-120,
    // whose integrate and meshing are commented out)
    {
        char text[1024];
        std::snprintf(text, sizeof text,
                      "%%YAML:1.0\n"
                      "tsdf.volume.voxel_grid_size.x: 60\ntsdf.volume.voxel_grid_size.y: 52\n"
                      "tsdf.volume.voxel_grid_size.z: 44\ntsdf.volume.physical_size.x: 0.9\n"
                      "tsdf.volume.physical_size.y: 0.78\ntsdf.volume.physical_size.z: 0.66\n"
                      "tsdf.volume.global_offset.x: -0.45\ntsdf.volume.global_offset.y: -0.39\n"
                      "tsdf.volume.global_offset.z: -0.33\ntsdf.volume.max_weight: 3\n"
                      "tsdf.volume.truncation: 0.04\n"
                      "camera.fx: %.17g\ncamera.fy: %.17g\ncamera.cx: %.17g\ncamera.cy: %.17g\n"
                      "camera.width: %d\ncamera.height: %d\n",
                      fx, fy, ccx, ccy, cols, rows);
        cuphoto::Config scfg;
        CHECK(scfg.parse(text), "parse");
        cuphoto::TSDFSurface surface(scfg);
        surface.set_global_transformation(poses[3]);
        CHECK(surface.reconstruct_surface(pc), "reconstruct_surface(cloud)");
        CHECK(!surface.reconstruct_surface(cuphoto::cudaPointCloud::Ptr()),
              "a null cloud must fail");
        const cuphoto::Mesh& mesh = surface.get_mesh();
        // oracle: the contiguous render, one integrate, marching cubes
        std::vector<float> cdepth(n1);
        std::vector<uint8_t> ccolor(3 * n1);
        oracle_render_points(all_xyz.data(), all_valid.data(), all_rgb.data(), (int64_t)np,
                             poses[3].data(), (float)fx, (float)fy, (float)ccx, (float)ccy, rows,
                             cols, cols, cdepth.data(), ccolor.data(), 3 * (int64_t)cols);
        std::vector<float> Ds(nvox), Ws(nvox);
        oracle_tsdf_init(&ov, Ds.data(), Ws.data());
        oracle_tsdf_integrate_f32(&ov, Ds.data(), Ws.data(), cdepth.data(), rows, cols, cols,
                                  poses[3].data(), nullptr, 0);
        const int64_t ns = oracle_mc_count(Ds.data(), ov.grid, nullptr, 0);
        std::vector<float> svs(9 * (size_t)ns + 9);
        std::vector<int32_t> sts(3 * (size_t)ns + 3);
        oracle_mc_extract(Ds.data(), ov.grid, ov.voxel_size, ov.offset, svs.data(), sts.data(),
                          nullptr, nullptr, ns, 0);
        CHECK((int64_t)mesh.polygons.size() == ns && ns > 100, "cloud mesh %zu vs %lld triangles",
              mesh.polygons.size(), (long long)ns);
        CHECK(mesh.vertices.size() == 3 * (size_t)ns &&
                  std::memcmp(mesh.vertices.data(), svs.data(), 36 * (size_t)ns) == 0,
              "cloud mesh vertices differ from the oracle");
    }

    (void)hipFree(ddepth);
    if (g_fail) {
        std::fprintf(stderr, "%d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("wrapper ok: %lld voxel updates, %zu triangles, bit-exact vs oracle\n",
                (long long)updates, tris.size());
    return 0;
}
