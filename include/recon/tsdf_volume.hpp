// tsdf_volume.hpp -- C++ drop-in for the TSDF API of cu-Reconstruction,
// layered on the C-ABI of recon.h (librecon_host.so over librecon.so).
//
// Same names, argument meanings and error behaviour as the reference:
//   TSDFVolumeConfig             modules/include/cuda/tsdf_volume.cuh:17-29
//   TSDFVolume                   modules/include/cuda/tsdf_volume.cuh:31-105
//   generate_triangular_surface  modules/include/cuda/marching_cubes.cuh:13-15
//   cudaPointCloud               modules/include/cuda/cuda_point_cloud.cuh:14-88
//                                (create, extract_points, project_to_depth,
//                                project_to_color_depth, get_points, ...)
//   write_mesh_to_ply            modules/src/utils.cpp:96-118
// Calls block until the device work is done (the reference synchronises the
// device after every kernel) and report failure as `false` plus a log line on
// stderr.  Images are anything shaped like cv::cuda::PtrStepSz<T> (members
// data, step in bytes, rows, cols), so OpenCV's types pass through unchanged.
// Differences that follow the build contract: colour is optional, the weight
// cap is applied, the sdf uses camera-space z, marching cubes uses the gx*gy
// layer stride, and voxel centres are computed, not stored
// (transformation_voxels_data() synthesises the reference's records on its
// first call).
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "recon/recon.h"

namespace cuphoto {

using r32 = float;
using r64 = double;
using ui32 = uint32_t;

// tsdf_volume.cuh:12-15 (declared, unused there): the observation weight of
// the integrate.  WEIGHT_BY_DEPTH weighs every observation 1 (the reference's
// update); WEIGHT_BY_VARIANCE weighs a depth sample d by sigma(1 m)^2 /
// sigma(d)^2, sigma(d) = a + b*(d - c)^2 (noise_model; 0 -> 0.0012, 0.0019, 0.4).
enum WEIGHT_TYPE {
    WEIGHT_BY_DEPTH = 0,
    WEIGHT_BY_VARIANCE,
};

struct TSDFVolumeConfig {
    dim3 voxel_grid_size;
    float3 physical_size;   // in physical units
    float3 global_offset;   // grid offset
    r32 max_weight;
    r32 focal_x;
    r32 focal_y;
    r32 cx;
    r32 cy;
    ui32 camera_width;
    ui32 camera_height;
    r32 max_cell_size = 0.0f;
    // additions (0 keeps the reference's constants 1.1*|voxel_size| and 24 m)
    r32 truncation = 0.0f;
    r32 max_depth = 0.0f;
    WEIGHT_TYPE weight_type = WEIGHT_BY_DEPTH;
    r32 noise_model[3] = {0.0f, 0.0f, 0.0f};
};

// Device image view with the members of cv::cuda::PtrStepSz<T>.
template <typename T>
struct PtrStepSz {
    T* data = nullptr;
    size_t step = 0;
    int rows = 0;
    int cols = 0;
    explicit operator bool() const { return data != nullptr; }
};
using PtrStepSzf = PtrStepSz<float>;
using PtrStepSzb = PtrStepSz<unsigned char>;

class TSDFVolume {
public:
    using Ptr = std::shared_ptr<TSDFVolume>;

    explicit TSDFVolume(const TSDFVolumeConfig& cfg, hipStream_t stream = nullptr);
    ~TSDFVolume();
    TSDFVolume(const TSDFVolume&) = delete;
    TSDFVolume& operator=(const TSDFVolume&) = delete;

    // tsdf_volume.cu:257-290; `color` is ignored, as in the reference, unless
    // enable_color() was called (then band voxels average it).  A colour
    // image given with data must cover the depth image (rows and cols in
    // pixels, as cv::cuda::PtrStepSzb of a CV_8UC3 GpuMat counts them), or
    // the call fails, as the reference fails on a zero-sized one (:265-268).
    template <typename Color, typename Depth>
    bool integrate(const Color& color, const Depth& depth, const std::array<r64, 7>& camera_pose) {
        if (color.data && (color.rows < depth.rows || color.cols < depth.cols)) {
            std::fprintf(stderr, "[cuphoto] TSDFVolume::integrate() -- color map (%dx%d) is smaller "
                                 "than the depth map (%dx%d)\n", (int)color.cols, (int)color.rows,
                         (int)depth.cols, (int)depth.rows);
            return false;
        }
        return integrate_raw(reinterpret_cast<const uint8_t*>(color.data), color.step,
                             reinterpret_cast<const float*>(depth.data), depth.step, depth.rows,
                             depth.cols, camera_pose);
    }
    // Frames f = 0..n-1 of a (n, rows, cols) stack, applied in order.
    bool integrate_frames(const float* depth0, size_t step, int rows, int cols,
                          size_t frame_stride_bytes, const std::vector<std::array<r64, 7>>& poses);

    r32* voxel_distances_data() const;
    r32* voxel_weights_data() const;
    dim3 voxel_grid_size() const { return cfg_.voxel_grid_size; }
    float3 global_offset() const { return cfg_.global_offset; }
    float3 scale_voxel_size() const;
    float3 voxel_phys_size() const { return cfg_.physical_size; }
    r32 truncation() const;

    // tsdf_volume.cuh:33-38: packed voxel transformation record
    struct TransformationVoxel {
        using Ptr = TransformationVoxel*;
        // plain members: HIP's float3/float4 are not POD, so a packed
        // struct of them would not be 28 bytes; .x/.y/.z(/.w) access is the same
        struct { float x, y, z; } translation;
        struct { float x, y, z, w; } quat_rotation;
    } __attribute__((packed));
    static_assert(sizeof(TransformationVoxel) == 28, "packed 28-byte record");
    // tsdf_volume.cuh:84-92.  Not stored by this library (centres are
    // computed); synthesised on the first call (28 B per voxel of device
    // memory, kept until the volume is destroyed), then returned as is.
    TransformationVoxel::Ptr transformation_voxels_data() const;

    bool ok() const { return h_ != nullptr; }
    void reset();
    // No counterpart in the reference's dense volume: a colour average per
    // voxel (r, g, b, weight) fed by integrate() calls with a colour image.
    bool enable_color(bool enable = true);
    // Colour (rgb bytes) at n device points (3 floats each) into device memory.
    bool sample_colors(const float* points, size_t n, uint8_t* colors) const;
    recon_tsdf* handle() const { return h_; }
    hipStream_t stream() const { return stream_; }

private:
    bool integrate_raw(const uint8_t* rgb, size_t rgb_step, const float* depth, size_t step,
                       int rows, int cols, const std::array<r64, 7>& pose);

    TSDFVolumeConfig cfg_;
    recon_tsdf* h_ = nullptr;
    hipStream_t stream_ = nullptr;
    mutable TransformationVoxel* t_voxels_ = nullptr;
};

// marching_cubes.cu:191-236: appends the triangle soup of the volume's zero
// level set (three fresh vertices per triangle, cube order z, y, x, slot).
void generate_triangular_surface(const TSDFVolume& tsdf_volume, std::vector<float3>& verts,
                                 std::vector<int3>& triangles);

// No counterpart in the reference (whose soup never shares a vertex): appends
// the welded mesh of the same surface -- one vertex per sign-changing grid
// edge, its unit normal from the TSDF gradient, and the soup's triangles as
// indices into `verts` (offset by the vertices already there).  `normals`
// grows with `verts`.
void generate_welded_surface(const TSDFVolume& tsdf_volume, std::vector<float3>& verts,
                             std::vector<float3>& normals, std::vector<int3>& triangles);

// cudaPointCloud (cuda_point_cloud.cuh:14-88): the device point buffer that
// extract_points fills and project_to_depth / project_to_color_depth render.
// Points are kept as three device planes -- xyz float3, a validity byte and
// an rgb triple -- which the kernels read and write coalesced; get_points()
// packs them into the reference's 33-byte Vertex records on request.
class cudaPointCloud {
public:
    // cuda_point_cloud.cuh:17-28: packed {float3 pos; float4 normal; uchar3
    // color; ui16 classID}.  Plain members: HIP's vector types are not POD,
    // so a packed struct of them would not be 33 bytes; .x/.y/.z access is
    // the same.
    struct Vertex {
        using Ptr = Vertex*;
        struct { float x, y, z; } pos;
        struct { float x, y, z, w; } normal;
        struct { uint8_t x, y, z; } color;
        uint16_t classID;
    } __attribute__((packed));
    static_assert(sizeof(Vertex) == 33, "packed 33-byte vertex");

    using Ptr = std::shared_ptr<cudaPointCloud>;

    // cuda_point_cloud.cu:121-124: camera_mat is the column-major 3x3
    // intrinsics (K[0], K[4], K[6], K[7] are fx, fy, cx, cy;
    // cuda_point_cloud.cu:44-47).  Throws std::bad_alloc when the device
    // buffers cannot be allocated (cuda_point_cloud.cu:107-112).
    static Ptr create(const std::array<r64, 9>& camera_mat, size_t total_number_pts) {
        Ptr p = std::make_shared<cudaPointCloud>(camera_mat, total_number_pts);
        if (p->get_total_num_points() != total_number_pts) throw std::bad_alloc();
        return p;
    }
    // As create(), with the stream the calls run on; allocation failure is
    // logged and leaves get_total_num_points() == 0.
    cudaPointCloud(const std::array<r64, 9>& camera_mat, size_t total_number_pts,
                   hipStream_t stream = nullptr);
    ~cudaPointCloud();
    cudaPointCloud(const cudaPointCloud&) = delete;
    cudaPointCloud& operator=(const cudaPointCloud&) = delete;

    // cuda_point_cloud.cu:179-226: back-projects frame `frame_idx` into
    // points [frame_idx*W*H, (frame_idx+1)*W*H).
    template <typename Depth, typename Color>
    bool extract_points(const Depth& depth, const Color& colors, const std::array<r64, 7>& transform,
                        int frame_idx) {
        return extract_raw(reinterpret_cast<const float*>(depth.data), depth.step, depth.rows,
                           depth.cols, reinterpret_cast<const uint8_t*>(colors.data), colors.step,
                           transform, frame_idx);
    }
    // cuda_point_cloud.cu:229-353 with filter_image_if (:89-99): z-buffer
    // render of the first get_num_points() points (invalid ones skipped) at
    // camera_pose (world->camera) into a pitched float depth map and, for
    // project_to_color_depth, a pitched rgb image.  A pixel keeps the nearest
    // point (the first of equal depths) and its colour; empty pixels get 0.
    // Runs on the device; blocks until done, as the reference does.  cx, cy
    // come from K[6], K[7]: the reference reads K[2], K[5], which are 0 in the
    // column-major K it fills (:248-251).
    template <typename Depth>
    bool project_to_depth(Depth& depth, const std::array<r64, 7>& camera_pose) {
        return render_raw(reinterpret_cast<float*>(depth.data), depth.step, depth.rows, depth.cols,
                          nullptr, 0, camera_pose, "cudaPointCloud::project_to_depth()");
    }
    template <typename Color, typename Depth>
    bool project_to_color_depth(Color& color, Depth& depth, const std::array<r64, 7>& camera_pose) {
        if (!color.data) {
            std::fprintf(stderr, "[cuphoto] cudaPointCloud::project_to_color_depth() -- color map "
                                 "is null pointer\n");
            return false;
        }
        return render_raw(reinterpret_cast<float*>(depth.data), depth.step, depth.rows, depth.cols,
                          reinterpret_cast<uint8_t*>(color.data), color.step, camera_pose,
                          "cudaPointCloud::project_to_color_depth()");
    }
    // cuda_point_cloud.cuh:40-53
    size_t get_total_num_points() const { return total_; }
    size_t get_num_points() const { return num_; }
    size_t get_size() const { return num_ * sizeof(Vertex); }
    std::array<r64, 9> get_camera_parameters() const { return K_; }
    void set_total_number_pts() { num_ = total_; dirty_ = true; }
    // The first get_total_num_points() points as Vertex records (normal 0,
    // classID 0; an invalid pixel is pos (0,0,0), colour (0,0,0)) in pinned
    // host memory that the host and the device both read, as the reference's
    // managed buffer (cuda_point_cloud.cu:170).  Packed again only after
    // extract_points, clear or set_total_number_pts changed the points;
    // nullptr on failure (logged).
    Vertex* get_points() const;
    Vertex* get_vertex(size_t index) const {
        Vertex* p = get_points();
        return p ? p + index : nullptr;
    }
    // cuda_point_cloud.cu:142-147: zeroes the points, num_points = 0.
    void clear();
    // copies the valid points to the host and writes an ASCII PLY
    // (cuda_point_cloud.cu:459-495, including its y/z sign flip)
    bool save_ply(const std::string& path) const;

    // The planes (device memory, get_total_num_points() entries each).
    float* xyz() const { return xyz_; }
    uint8_t* valid() const { return valid_; }
    uint8_t* colors() const { return rgb_; }

private:
    bool extract_raw(const float* depth, size_t step, int rows, int cols, const uint8_t* rgb,
                     size_t rgb_step, const std::array<r64, 7>& T, int frame_idx);
    bool render_raw(float* depth, size_t step, int rows, int cols, uint8_t* color,
                    size_t color_step, const std::array<r64, 7>& pose, const char* who);
    std::array<r64, 9> K_;
    unsigned long long* zbuffer_ = nullptr;  // grows to rows*cols words
    size_t zbuffer_words_ = 0;
    size_t total_ = 0, num_ = 0;
    float* xyz_ = nullptr;
    uint8_t* valid_ = nullptr;
    uint8_t* rgb_ = nullptr;
    hipStream_t stream_ = nullptr;
    mutable Vertex* packed_ = nullptr;  // pinned host memory, total_ records
    mutable bool dirty_ = true;
};
// The name this library used for the same class.
using PointCloud = cudaPointCloud;

// utils.cpp:96-118: ASCII PLY of vertices and triangles.
bool write_mesh_to_ply(const std::string& file_name, const std::vector<float3>& vertices,
                       const std::vector<int3>& triangles);
// The same with per-vertex normals (property float nx, ny, nz after x, y, z).
bool write_mesh_to_ply(const std::string& file_name, const std::vector<float3>& vertices,
                       const std::vector<float3>& normals, const std::vector<int3>& triangles);
// The same with optional per-vertex normals and colours (either vector may be
// empty; property uchar red, green, blue after the normals).
bool write_mesh_to_ply(const std::string& file_name, const std::vector<float3>& vertices,
                       const std::vector<float3>& normals, const std::vector<uchar3>& colors,
                       const std::vector<int3>& triangles);

}  // namespace cuphoto
