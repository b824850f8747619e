// tsdf_surface.cpp -- TSDFSurface, the RGB-D directory dataset and the mesh
// store of the reference's TSDF driver
// (modules/src/tsdf_surface_reconstruction.cpp:13-153,
//  examples/common/rgbd_dataset.cpp:10-36, modules/src/utils.cpp:122-145,
//  modules/src/mesh.cpp:59-64) over the C++ drop-in of this library.
#include "recon/tsdf_surface.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <filesystem>

namespace cuphoto {

std::array<double, 7> euler_pose(double roll_deg, double yaw_deg, double pitch_deg, double tx,
                                 double ty, double tz) {
    // angles as the reference forms them: float degrees times a float
    // degree-to-radian constant (tsdf_surface_reconstruction.cpp:70-73)
    const float half_c = (float)(M_PI / 180.);
    const double roll = (double)((float)roll_deg * half_c);
    const double yaw = (double)((float)yaw_deg * half_c);
    const double pitch = (double)((float)pitch_deg * half_c);
    // q = qz(roll) * qy(yaw) * qx(pitch), Hamilton products
    const double qz[4] = {std::cos(roll / 2), 0.0, 0.0, std::sin(roll / 2)};
    const double qy[4] = {std::cos(yaw / 2), 0.0, std::sin(yaw / 2), 0.0};
    const double qx[4] = {std::cos(pitch / 2), std::sin(pitch / 2), 0.0, 0.0};
    auto mul = [](const double* a, const double* b, double* o) {
        o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
        o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
        o[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
        o[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    };
    double t1[4], q[4];
    mul(qz, qy, t1);
    mul(t1, qx, q);
    return {q[0], q[1], q[2], q[3], (double)(float)tx, (double)(float)ty, (double)(float)tz};
}

TSDFSurface::TSDFSurface(const Config& cfg) {
    cfg_.voxel_grid_size.x = (unsigned)cfg.get<float>("tsdf.volume.voxel_grid_size.x");
    cfg_.voxel_grid_size.y = (unsigned)cfg.get<float>("tsdf.volume.voxel_grid_size.y");
    cfg_.voxel_grid_size.z = (unsigned)cfg.get<float>("tsdf.volume.voxel_grid_size.z");
    cfg_.physical_size.x = cfg.get<float>("tsdf.volume.physical_size.x");
    cfg_.physical_size.y = cfg.get<float>("tsdf.volume.physical_size.y");
    cfg_.physical_size.z = cfg.get<float>("tsdf.volume.physical_size.z");
    cfg_.max_weight = cfg.get<float>("tsdf.volume.max_weight");
    cfg_.global_offset.x = cfg.get<float>("tsdf.volume.global_offset.x", 0.f);
    cfg_.global_offset.y = cfg.get<float>("tsdf.volume.global_offset.y", 0.f);
    cfg_.global_offset.z = cfg.get<float>("tsdf.volume.global_offset.z", 0.f);
    cfg_.focal_x = cfg.get<float>("camera.fx", 1.f);
    cfg_.focal_y = cfg.get<float>("camera.fy", 1.f);
    cfg_.cx = cfg.get<float>("camera.cx", 0.f);
    cfg_.cy = cfg.get<float>("camera.cy", 0.f);
    cfg_.camera_width = (ui32)cfg.get<float>("camera.width");
    cfg_.camera_height = (ui32)cfg.get<float>("camera.height");
    // additions of this library (0: the reference's constants)
    cfg_.truncation = cfg.get<float>("tsdf.volume.truncation", 0.f);
    cfg_.max_depth = cfg.get<float>("tsdf.volume.max_depth", 0.f);
    cfg_.weight_type = cfg.get<int>("tsdf.volume.weight_type", 0) == 1 ? WEIGHT_BY_VARIANCE : WEIGHT_BY_DEPTH;
    welded_ = cfg.get<int>("tsdf.mesh.welded", 0) != 0;
    color_ = cfg.get<int>("dataset.color.mode", 0) != 0;
    global_pose = euler_pose(cfg.get<float>("tsdf.volume.camera.pose.euler.roll", 0.f),
                             cfg.get<float>("tsdf.volume.camera.pose.euler.yaw", 0.f),
                             cfg.get<float>("tsdf.volume.camera.pose.euler.pitch", 0.f),
                             cfg.get<float>("tsdf.volume.camera.pose.translation.x", 0.f),
                             cfg.get<float>("tsdf.volume.camera.pose.translation.y", 0.f),
                             cfg.get<float>("tsdf.volume.camera.pose.translation.z", 0.f));
}

TSDFVolumeConfig TSDFSurface::build_cfg() const { return cfg_; }

bool TSDFSurface::reconstruct(const PtrStepSz<float>& depth, const unsigned char* rgb,
                              size_t rgb_step) {
    TSDFVolume volume(cfg_);
    if (!volume.ok()) return false;
    const bool colored = color_ && rgb;
    if (colored && !volume.enable_color()) return false;
    PtrStepSz<unsigned char> color{const_cast<unsigned char*>(rgb), rgb_step, depth.rows, depth.cols};
    if (!volume.integrate(color, depth, global_pose)) return false;
    std::vector<float3> verts, normals;
    std::vector<int3> triangles;
    if (welded_)
        generate_welded_surface(volume, verts, normals, triangles);
    else
        generate_triangular_surface(volume, verts, triangles);
    std::vector<uchar3> colors;
    if (colored && !verts.empty()) {
        // vertex colours from the colour volume (recon_tsdf_sample_colors)
        float* dp = nullptr;
        uint8_t* dc = nullptr;
        bool ok = hipMalloc((void**)&dp, 12 * verts.size()) == hipSuccess &&
                  hipMalloc((void**)&dc, 3 * verts.size()) == hipSuccess &&
                  hipMemcpy(dp, verts.data(), 12 * verts.size(), hipMemcpyHostToDevice) == hipSuccess &&
                  volume.sample_colors(dp, verts.size(), dc);
        std::vector<uint8_t> host(3 * verts.size());
        ok = ok && hipMemcpy(host.data(), dc, host.size(), hipMemcpyDeviceToHost) == hipSuccess;
        if (dp) (void)hipFree(dp);
        if (dc) (void)hipFree(dc);
        if (!ok) {
            std::fprintf(stderr, "[cuphoto] TSDFSurface::reconstruct_surface() -- vertex colours failed\n");
            return false;
        }
        colors.resize(verts.size());
        for (size_t i = 0; i < verts.size(); ++i)
            colors[i] = make_uchar3(host[3 * i], host[3 * i + 1], host[3 * i + 2]);
    }
    surface_mesh.vertices = std::move(verts);
    surface_mesh.normals = std::move(normals);
    surface_mesh.colors = std::move(colors);
    surface_mesh.polygons = std::move(triangles);
    return true;
}

bool TSDFSurface::reconstruct_surface(const cudaPointCloud::Ptr cuda_pc) {
    if (!cuda_pc) {
        std::fprintf(stderr, "[cuphoto] TSDFSurface::reconstruct_surface() -- point cloud is null\n");
        return false;
    }
    return reconstruct_surface(*cuda_pc);
}

bool TSDFSurface::reconstruct_surface(cudaPointCloud& cloud) {
    const int rows = (int)cfg_.camera_height, cols = (int)cfg_.camera_width;
    if (rows <= 0 || cols <= 0) {
        std::fprintf(stderr, "[cuphoto] TSDFSurface::reconstruct_surface() -- camera size is zero\n");
        return false;
    }
    float* depth = nullptr;
    unsigned char* color = nullptr;
    const size_t n = (size_t)rows * cols;
    bool ok = hipMalloc((void**)&depth, 4 * n) == hipSuccess &&
              hipMalloc((void**)&color, 3 * n) == hipSuccess;
    if (!ok) std::fprintf(stderr, "[cuphoto] TSDFSurface::reconstruct_surface() -- out of device memory\n");
    PtrStepSz<float> d{depth, 4 * (size_t)cols, rows, cols};
    PtrStepSz<unsigned char> c{color, 3 * (size_t)cols, rows, cols};
    ok = ok && cloud.project_to_color_depth(c, d, global_pose) && reconstruct(d, color, 3 * (size_t)cols);
    if (depth) (void)hipFree(depth);
    if (color) (void)hipFree(color);
    return ok;
}

bool Mesh::store_to_ply(const std::string& ply_filepath) const {
    if (!normals.empty() || !colors.empty())
        return write_mesh_to_ply(ply_filepath, vertices, normals, colors, polygons);
    return write_mesh_to_ply(ply_filepath, vertices, polygons);
}

std::set<std::string> files_directory(const std::string& data_path) {
    static const std::set<std::string> extensions{".jpg", ".jpeg", ".png", ".tiff"};
    std::set<std::string> files;
    std::error_code ec;
    for (const auto& entry : std::filesystem::directory_iterator(
             data_path, std::filesystem::directory_options::skip_permission_denied, ec)) {
        if (!extensions.count(entry.path().extension().string())) continue;
        const auto type = entry.status(ec).type();
        if (type != std::filesystem::file_type::regular && type != std::filesystem::file_type::symlink)
            continue;
        files.insert(std::filesystem::canonical(entry.path(), ec).string());
    }
    return files;
}

RGBDDataset::RGBDDataset(const std::string& data_path, double scale) : depth_scale(scale) {
    const auto set = files_directory(data_path);
    files.assign(set.begin(), set.end());
}

RGBDDataset::~RGBDDataset() {
    if (device_depth) (void)hipFree(device_depth);
    if (device_color) (void)hipFree(device_color);
}

bool RGBDDataset::get_next(PtrStepSz<float>* depth, PtrStepSz<unsigned char>* color) {
    if (next + 1 >= files.size()) {
        std::fprintf(stderr, "[cuphoto] all files in the directory were traversed\n");
        return false;
    }
    const std::string& color_file = files[next];
    const std::string& depth_file = files[next + 1];  // the colour file comes first
    next += 2;
    if (color) {
        int cw = 0, ch = 0;
        std::vector<uint8_t> rgb;
        std::string cerr;
        if (!read_rgb_png(color_file, &cw, &ch, &rgb, &cerr)) {
            std::fprintf(stderr, "[cuphoto] %s\n", cerr.c_str());
            return false;
        }
        if (rgb.size() > device_color_bytes) {
            if (device_color) (void)hipFree(device_color);
            device_color = nullptr;
            device_color_bytes = 0;
            if (hipMalloc((void**)&device_color, rgb.size()) != hipSuccess) {
                std::fprintf(stderr, "[cuphoto] out of device memory for a colour frame\n");
                return false;
            }
            device_color_bytes = rgb.size();
        }
        if (hipMemcpy(device_color, rgb.data(), rgb.size(), hipMemcpyHostToDevice) != hipSuccess) {
            std::fprintf(stderr, "[cuphoto] colour upload failed\n");
            return false;
        }
        *color = PtrStepSz<unsigned char>{device_color, 3 * (size_t)cw, ch, cw};
    }
    GrayImage img;
    std::string err;
    if (!read_gray_png(depth_file, &img, &err)) {
        std::fprintf(stderr, "[cuphoto] %s\n", err.c_str());
        return false;
    }
    // depth in metres: raw * depth_scale, rounded once to float
    std::vector<float> host((size_t)img.width * img.height);
    for (size_t i = 0; i < host.size(); ++i) host[i] = (float)((double)img.pixels[i] * depth_scale);
    const size_t bytes = host.size() * sizeof(float);
    if (bytes > device_bytes) {
        if (device_depth) (void)hipFree(device_depth);
        device_depth = nullptr;
        device_bytes = 0;
        if (hipMalloc((void**)&device_depth, bytes) != hipSuccess) {
            std::fprintf(stderr, "[cuphoto] out of device memory for a depth frame\n");
            return false;
        }
        device_bytes = bytes;
    }
    if (hipMemcpy(device_depth, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        std::fprintf(stderr, "[cuphoto] depth upload failed\n");
        return false;
    }
    *depth = PtrStepSz<float>{device_depth, (size_t)img.width * sizeof(float), img.height, img.width};
    return true;
}

}  // namespace cuphoto
