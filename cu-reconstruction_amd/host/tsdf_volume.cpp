// tsdf_volume.cpp -- the C++ drop-in of include/recon/tsdf_volume.hpp over the
// C-ABI of recon.h.  Every call waits for its stream, as the reference's
// calls wait for the device (tsdf_volume.cu:253,287; marching_cubes.cu:223).
#include "recon/tsdf_volume.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <fstream>

namespace cuphoto {

namespace {

void log_error(const char* where) {
    std::fprintf(stderr, "[cuphoto] %s -- %s\n", where, recon_last_error());
}

recon_tsdf_config to_c(const TSDFVolumeConfig& c) {
    recon_tsdf_config r{};
    r.voxel_grid_size[0] = (int32_t)c.voxel_grid_size.x;
    r.voxel_grid_size[1] = (int32_t)c.voxel_grid_size.y;
    r.voxel_grid_size[2] = (int32_t)c.voxel_grid_size.z;
    r.physical_size[0] = c.physical_size.x;
    r.physical_size[1] = c.physical_size.y;
    r.physical_size[2] = c.physical_size.z;
    r.global_offset[0] = c.global_offset.x;
    r.global_offset[1] = c.global_offset.y;
    r.global_offset[2] = c.global_offset.z;
    r.max_weight = c.max_weight;
    r.focal_x = c.focal_x;
    r.focal_y = c.focal_y;
    r.cx = c.cx;
    r.cy = c.cy;
    r.camera_width = c.camera_width;
    r.camera_height = c.camera_height;
    r.max_cell_size = c.max_cell_size;
    r.truncation = c.truncation;
    r.max_depth = c.max_depth;
    r.weight_type = (int32_t)c.weight_type;
    for (int i = 0; i < 3; ++i) r.noise_model[i] = c.noise_model[i];
    return r;
}

bool sync(hipStream_t s, const char* where) {
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        std::fprintf(stderr, "[cuphoto] %s -- %s\n", where, hipGetErrorString(e));
        return false;
    }
    return true;
}

}  // namespace

TSDFVolume::TSDFVolume(const TSDFVolumeConfig& cfg, hipStream_t stream) : cfg_(cfg), stream_(stream) {
    const recon_tsdf_config c = to_c(cfg);
    if (recon_tsdf_create(&c, &h_) != RECON_OK) {
        log_error("TSDFVolume::TSDFVolume()");
        h_ = nullptr;
    }
}

TSDFVolume::~TSDFVolume() {
    if (t_voxels_) (void)hipFree(t_voxels_);
    recon_tsdf_destroy(h_);
}

bool TSDFVolume::integrate_raw(const uint8_t* rgb, size_t rgb_step, const float* depth, size_t step,
                               int rows, int cols, const std::array<r64, 7>& pose) {
    if (!h_) {
        std::fprintf(stderr, "[cuphoto] TSDFVolume::integrate() -- volume was not created\n");
        return false;
    }
    const recon_depth_view dv{depth, step, rows, cols};
    if (recon_tsdf_integrate(h_, &dv, rgb, rgb_step, pose.data(), stream_) != RECON_OK) {
        log_error("TSDFVolume::integrate()");
        return false;
    }
    return sync(stream_, "TSDFVolume::integrate()");
}

TSDFVolume::TransformationVoxel::Ptr TSDFVolume::transformation_voxels_data() const {
    if (t_voxels_ || !h_) return t_voxels_;
    recon_tsdf_info info;
    if (recon_tsdf_get_info(h_, &info) != RECON_OK) return nullptr;
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(TransformationVoxel) * (size_t)info.num_voxels) != hipSuccess) {
        std::fprintf(stderr, "[cuphoto] TSDFVolume::transformation_voxels_data() -- out of device memory\n");
        return nullptr;
    }
    if (recon_tsdf_transformation_voxels(h_, p, stream_) != RECON_OK ||
        !sync(stream_, "TSDFVolume::transformation_voxels_data()")) {
        log_error("TSDFVolume::transformation_voxels_data()");
        (void)hipFree(p);
        return nullptr;
    }
    t_voxels_ = static_cast<TransformationVoxel*>(p);
    return t_voxels_;
}

bool TSDFVolume::enable_color(bool enable) {
    if (!h_ || recon_tsdf_enable_color(h_, enable ? 1 : 0, stream_) != RECON_OK) {
        log_error("TSDFVolume::enable_color()");
        return false;
    }
    return sync(stream_, "TSDFVolume::enable_color()");
}

bool TSDFVolume::sample_colors(const float* points, size_t n, uint8_t* colors) const {
    if (!h_ || recon_tsdf_sample_colors(h_, points, (int64_t)n, colors, stream_) != RECON_OK) {
        log_error("TSDFVolume::sample_colors()");
        return false;
    }
    return sync(stream_, "TSDFVolume::sample_colors()");
}

bool TSDFVolume::integrate_frames(const float* depth0, size_t step, int rows, int cols,
                                  size_t frame_stride_bytes,
                                  const std::vector<std::array<r64, 7>>& poses) {
    if (!h_) return false;
    const recon_depth_view dv{depth0, step, rows, cols};
    std::vector<double> flat;
    flat.reserve(7 * poses.size());
    for (const auto& p : poses) flat.insert(flat.end(), p.begin(), p.end());
    if (recon_tsdf_integrate_frames(h_, &dv, frame_stride_bytes, (int32_t)poses.size(), flat.data(),
                                    stream_) != RECON_OK) {
        log_error("TSDFVolume::integrate_frames()");
        return false;
    }
    return sync(stream_, "TSDFVolume::integrate_frames()");
}

r32* TSDFVolume::voxel_distances_data() const { return recon_tsdf_distances(h_); }
r32* TSDFVolume::voxel_weights_data() const { return recon_tsdf_weights(h_); }

float3 TSDFVolume::scale_voxel_size() const {
    recon_tsdf_info info{};
    if (!h_ || recon_tsdf_get_info(h_, &info) != RECON_OK) return make_float3(0.f, 0.f, 0.f);
    return make_float3(info.voxel_size[0], info.voxel_size[1], info.voxel_size[2]);
}

r32 TSDFVolume::truncation() const {
    recon_tsdf_info info{};
    if (!h_ || recon_tsdf_get_info(h_, &info) != RECON_OK) return 0.f;
    return info.truncation;
}

void TSDFVolume::reset() {
    if (h_ && recon_tsdf_reset(h_, stream_) == RECON_OK) sync(stream_, "TSDFVolume::reset()");
}

void generate_triangular_surface(const TSDFVolume& tsdf_volume, std::vector<float3>& verts,
                                 std::vector<int3>& triangles) {
    recon_tsdf* h = tsdf_volume.handle();
    hipStream_t s = tsdf_volume.stream();
    int64_t n = 0;
    if (!h || recon_tsdf_count_triangles(h, &n, s) != RECON_OK) {
        log_error("generate_triangular_surface()");
        return;
    }
    if (n == 0) return;
    float* dverts = nullptr;
    if (hipMalloc((void**)&dverts, sizeof(float) * 9 * (size_t)n) != hipSuccess) {
        std::fprintf(stderr, "[cuphoto] generate_triangular_surface() -- out of device memory\n");
        return;
    }
    int64_t got = 0;
    if (recon_tsdf_extract_mesh(h, dverts, nullptr, n, &got, s) != RECON_OK) {
        log_error("generate_triangular_surface()");
        (void)hipFree(dverts);
        return;
    }
    // append, as the reference's host loop does (marching_cubes.cu:176-183):
    // triangle t refers to the three vertices it appended
    const size_t v0 = verts.size();
    verts.resize(v0 + 3 * (size_t)got);
    const hipError_t e = hipMemcpy(verts.data() + v0, dverts, sizeof(float) * 9 * (size_t)got,
                                   hipMemcpyDeviceToHost);
    (void)hipFree(dverts);
    if (e != hipSuccess) {
        verts.resize(v0);
        std::fprintf(stderr, "[cuphoto] generate_triangular_surface() -- %s\n", hipGetErrorString(e));
        return;
    }
    triangles.reserve(triangles.size() + (size_t)got);
    for (int64_t t = 0; t < got; ++t) {
        const int b = (int)(v0 + 3 * t);
        triangles.push_back(make_int3(b, b + 1, b + 2));
    }
}

cudaPointCloud::cudaPointCloud(const std::array<r64, 9>& camera_mat, size_t total_number_pts,
                               hipStream_t stream)
    : K_(camera_mat), total_(total_number_pts), stream_(stream) {
    if (total_ == 0) return;
    if (hipMalloc((void**)&xyz_, 12 * total_) != hipSuccess ||
        hipMalloc((void**)&valid_, total_) != hipSuccess ||
        hipMalloc((void**)&rgb_, 3 * total_) != hipSuccess) {
        std::fprintf(stderr, "[cuphoto] cudaPointCloud::reserve_memory() -- can't allocate %zu points\n",
                     total_);
        total_ = 0;
    }
}

cudaPointCloud::~cudaPointCloud() {
    if (xyz_) (void)hipFree(xyz_);
    if (valid_) (void)hipFree(valid_);
    if (rgb_) (void)hipFree(rgb_);
    if (zbuffer_) (void)hipFree(zbuffer_);
    if (packed_) (void)hipHostFree(packed_);
}

bool cudaPointCloud::render_raw(float* depth, size_t step, int rows, int cols, uint8_t* color,
                                size_t color_step, const std::array<r64, 7>& pose, const char* who) {
    if (!depth) {  // cuda_point_cloud.cu:231-234
        std::fprintf(stderr, "[cuphoto] %s -- depth map is null pointer\n", who);
        return false;
    }
    if (rows <= 0 || cols <= 0) {  // :239-242
        std::fprintf(stderr, "[cuphoto] %s -- depth width/height parameters are zero\n", who);
        return false;
    }
    const size_t words = (size_t)rows * (size_t)cols;
    if (words > zbuffer_words_) {
        if (zbuffer_) (void)hipFree(zbuffer_);
        zbuffer_ = nullptr;
        zbuffer_words_ = 0;
        if (hipMalloc((void**)&zbuffer_, 8 * words) != hipSuccess) {
            std::fprintf(stderr, "[cuphoto] %s -- can't allocate the z-buffer\n", who);
            return false;
        }
        zbuffer_words_ = words;
    }
    if (recon_points_render(xyz_, valid_, rgb_, (int64_t)num_, pose.data(), K_[0], K_[4], K_[6],
                            K_[7], depth, step, rows, cols, color, color_step, zbuffer_,
                            stream_) != RECON_OK) {
        log_error(who);
        return false;
    }
    return sync(stream_, who);
}

bool cudaPointCloud::extract_raw(const float* depth, size_t step, int rows, int cols,
                                 const uint8_t* rgb, size_t rgb_step, const std::array<r64, 7>& T,
                                 int frame_idx) {
    if (!depth) {
        std::fprintf(stderr, "[cuphoto] cudaPointCloud::extract_points() -- depth map is null pointer\n");
        return false;
    }
    const size_t n = (size_t)rows * (size_t)cols;
    const size_t start = (size_t)frame_idx * n;
    if (frame_idx < 0 || start + n > total_) {
        std::fprintf(stderr, "[cuphoto] cudaPointCloud::extract_points() -- frame %d does not fit\n",
                     frame_idx);
        return false;
    }
    const recon_depth_view dv{depth, step, rows, cols};
    // column-major K: cx, cy at K[2][0], K[2][1] (cuda_point_cloud.cu:44-47)
    if (recon_depth_to_points(&dv, K_[0], K_[4], K_[6], K_[7], T.data(), rgb, rgb_step,
                              xyz_ + 3 * start, valid_ + start, rgb_ + 3 * start, stream_) != RECON_OK) {
        log_error("cudaPointCloud::extract_points()");
        return false;
    }
    num_ += n;  // cuda_point_cloud.cu:221 counts every pixel
    dirty_ = true;
    return sync(stream_, "cudaPointCloud::extract_points()");
}

cudaPointCloud::Vertex* cudaPointCloud::get_points() const {
    if (total_ == 0) return nullptr;
    if (!packed_) {
        void* p = nullptr;
        if (hipHostMalloc(&p, sizeof(Vertex) * total_, hipHostMallocDefault) != hipSuccess) {
            std::fprintf(stderr, "[cuphoto] cudaPointCloud::get_points() -- can't allocate %zu "
                                 "vertices\n", total_);
            return nullptr;
        }
        packed_ = static_cast<Vertex*>(p);
        dirty_ = true;
    }
    if (!dirty_) return packed_;
    std::vector<float> xyz(3 * total_);
    std::vector<uint8_t> rgb(3 * total_);
    if (!sync(stream_, "cudaPointCloud::get_points()") ||
        hipMemcpy(xyz.data(), xyz_, 12 * total_, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(rgb.data(), rgb_, 3 * total_, hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "[cuphoto] cudaPointCloud::get_points() -- copy failed\n");
        return nullptr;
    }
    for (size_t i = 0; i < total_; ++i) {
        Vertex& v = packed_[i];
        v.pos.x = xyz[3 * i];
        v.pos.y = xyz[3 * i + 1];
        v.pos.z = xyz[3 * i + 2];
        v.normal.x = v.normal.y = v.normal.z = v.normal.w = 0.f;
        v.color.x = rgb[3 * i];
        v.color.y = rgb[3 * i + 1];
        v.color.z = rgb[3 * i + 2];
        v.classID = 0;
    }
    dirty_ = false;
    return packed_;
}

void cudaPointCloud::clear() {
    if (total_ && (hipMemsetAsync(xyz_, 0, 12 * total_, stream_) != hipSuccess ||
                   hipMemsetAsync(valid_, 0, total_, stream_) != hipSuccess ||
                   hipMemsetAsync(rgb_, 0, 3 * total_, stream_) != hipSuccess))
        std::fprintf(stderr, "[cuphoto] cudaPointCloud::clear() -- memset failed\n");
    (void)sync(stream_, "cudaPointCloud::clear()");
    num_ = 0;
    dirty_ = true;
}

bool cudaPointCloud::save_ply(const std::string& path) const {
    if (path.empty() || num_ == 0) return false;
    std::vector<float> xyz(3 * num_);
    std::vector<uint8_t> valid(num_), rgb(3 * num_);
    if (hipMemcpy(xyz.data(), xyz_, 12 * num_, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(valid.data(), valid_, num_, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(rgb.data(), rgb_, 3 * num_, hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    size_t nv = 0;
    for (uint8_t v : valid) nv += v;
    std::ofstream of(path);
    if (!of) return false;
    of << "ply\nformat ascii 1.0\nelement vertex " << nv
       << "\nproperty float x\nproperty float y\nproperty float z\nproperty uchar red\n"
          "property uchar green\nproperty uchar blue\nend_header\n";
    for (size_t i = 0; i < num_; ++i) {
        if (!valid[i]) continue;
        of << xyz[3 * i] << " " << xyz[3 * i + 1] * -1.f << " " << xyz[3 * i + 2] * -1.f << " "
           << (int)rgb[3 * i] << " " << (int)rgb[3 * i + 1] << " " << (int)rgb[3 * i + 2] << "\n";
    }
    return (bool)of;
}

void generate_welded_surface(const TSDFVolume& tsdf_volume, std::vector<float3>& verts,
                             std::vector<float3>& normals, std::vector<int3>& triangles) {
    recon_tsdf* h = tsdf_volume.handle();
    hipStream_t s = tsdf_volume.stream();
    int64_t nv = 0, nt = 0;
    if (!h || recon_tsdf_extract_welded_mesh(h, nullptr, nullptr, 0, nullptr, 0, &nv, &nt, s) != RECON_OK) {
        log_error("generate_welded_surface()");
        return;
    }
    if (nv == 0 || nt == 0) return;
    float* dv = nullptr;
    float* dn = nullptr;
    int32_t* dt = nullptr;
    bool ok = hipMalloc((void**)&dv, 12 * (size_t)nv) == hipSuccess &&
              hipMalloc((void**)&dn, 12 * (size_t)nv) == hipSuccess &&
              hipMalloc((void**)&dt, 12 * (size_t)nt) == hipSuccess;
    if (!ok) std::fprintf(stderr, "[cuphoto] generate_welded_surface() -- out of device memory\n");
    ok = ok && recon_tsdf_extract_welded_mesh(h, dv, dn, nv, dt, nt, &nv, &nt, s) == RECON_OK;
    if (!ok) log_error("generate_welded_surface()");
    const size_t v0 = verts.size(), t0 = triangles.size();
    if (ok) {
        verts.resize(v0 + (size_t)nv);
        normals.resize(v0 + (size_t)nv);
        triangles.resize(t0 + (size_t)nt);
        ok = hipMemcpy(verts.data() + v0, dv, 12 * (size_t)nv, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(normals.data() + v0, dn, 12 * (size_t)nv, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(triangles.data() + t0, dt, 12 * (size_t)nt, hipMemcpyDeviceToHost) == hipSuccess;
        if (!ok) {
            std::fprintf(stderr, "[cuphoto] generate_welded_surface() -- copy failed\n");
            verts.resize(v0);
            normals.resize(v0);
            triangles.resize(t0);
        } else if (v0) {
            for (size_t t = t0; t < triangles.size(); ++t) {
                triangles[t].x += (int)v0;
                triangles[t].y += (int)v0;
                triangles[t].z += (int)v0;
            }
        }
    }
    if (dv) (void)hipFree(dv);
    if (dn) (void)hipFree(dn);
    if (dt) (void)hipFree(dt);
}

bool write_mesh_to_ply(const std::string& file_name, const std::vector<float3>& vertices,
                       const std::vector<int3>& triangles) {
    std::ofstream of(file_name);
    if (!of.is_open()) return false;
    of << "ply"
       << '\n' << "format ascii 1.0"
       << '\n' << "element vertex " << vertices.size()
       << '\n' << "property float x"
       << '\n' << "property float y"
       << '\n' << "property float z"
       << '\n' << "element face " << triangles.size()
       << '\n' << "property list uchar int vertex_indices"
       << '\n' << "end_header\n";
    for (const float3& v : vertices) of << v.x << " " << v.y << " " << v.z << "\n";
    for (const int3& t : triangles) of << "3 " << t.x << " " << t.y << " " << t.z << "\n";
    return (bool)of;
}

}  // namespace cuphoto

namespace cuphoto {

bool write_mesh_to_ply(const std::string& file_name, const std::vector<float3>& vertices,
                       const std::vector<float3>& normals, const std::vector<int3>& triangles) {
    if (normals.size() != vertices.size()) return false;
    return write_mesh_to_ply(file_name, vertices, normals, std::vector<uchar3>(), triangles);
}

bool write_mesh_to_ply(const std::string& file_name, const std::vector<float3>& vertices,
                       const std::vector<float3>& normals, const std::vector<uchar3>& colors,
                       const std::vector<int3>& triangles) {
    const bool hn = !normals.empty(), hc = !colors.empty();
    if ((hn && normals.size() != vertices.size()) || (hc && colors.size() != vertices.size()))
        return false;
    std::ofstream of(file_name);
    if (!of.is_open()) return false;
    of << "ply\nformat ascii 1.0\nelement vertex " << vertices.size()
       << "\nproperty float x\nproperty float y\nproperty float z";
    if (hn) of << "\nproperty float nx\nproperty float ny\nproperty float nz";
    if (hc) of << "\nproperty uchar red\nproperty uchar green\nproperty uchar blue";
    of << "\nelement face " << triangles.size()
       << "\nproperty list uchar int vertex_indices\nend_header\n";
    for (size_t i = 0; i < vertices.size(); ++i) {
        of << vertices[i].x << " " << vertices[i].y << " " << vertices[i].z;
        if (hn) of << " " << normals[i].x << " " << normals[i].y << " " << normals[i].z;
        if (hc) of << " " << (int)colors[i].x << " " << (int)colors[i].y << " " << (int)colors[i].z;
        of << "\n";
    }
    for (const int3& t : triangles) of << "3 " << t.x << " " << t.y << " " << t.z << "\n";
    return (bool)of;
}

}  // namespace cuphoto
