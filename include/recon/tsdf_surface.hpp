// tsdf_surface.hpp -- the reference's TSDF driver over this library: the
// parameter file, the RGB-D directory dataset, TSDFSurface and the mesh.
//
// Replaces (cu-Reconstruction paths):
//   cuphoto::Config                modules/include/config.hpp, modules/src/config.cpp
//                                  (an OpenCV FileStorage YAML of flat "key: value" lines)
//   cuphoto::RGBDDataset           examples/common/rgbd_dataset.cpp:10-36,
//                                  modules/src/utils.cpp:122-145 (sorted image files,
//                                  colour then depth; depth * depth_scale -> metres)
//   cuphoto::TSDFSurface           modules/src/tsdf_surface_reconstruction.cpp:13-153
//   cuphoto::Mesh::store_to_ply    modules/src/mesh.cpp:59-64, modules/src/utils.cpp:96-118
//   the tsdf_surface example       examples/tsdf_surface_reconstruction/tsdf_surface.cpp
//
// Images are decoded without OpenCV: the depth file must be a PNG (grey, 8 or
// 16 bit, the Kinect format), the colour file is only counted (integrate does
// not read colour).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "recon/tsdf_volume.hpp"

namespace cuphoto {

// Flat YAML parameter file: "%YAML:1.0" header, "key: value" lines, '#'
// comments, quoted or bare strings.  get<T>(key, default) returns `def` for
// a missing key, as FileStorage yields an empty node.
class Config {
public:
    bool set_parameter_file(const std::string& filename);
    bool parse(const std::string& text);
    bool has(const std::string& key) const { return values.count(key) != 0; }
    template <class T>
    T get(const std::string& key, T def = T()) const;

private:
    std::map<std::string, std::string> values;
};

template <> double Config::get<double>(const std::string& key, double def) const;
template <> float Config::get<float>(const std::string& key, float def) const;
template <> int Config::get<int>(const std::string& key, int def) const;
template <> std::string Config::get<std::string>(const std::string& key, std::string def) const;

// Decoded image (8- or 16-bit samples widened to uint16, channels
// interleaved).
struct GrayImage {
    int width = 0, height = 0, bit_depth = 0, channels = 1;
    std::vector<uint16_t> pixels;  // row-major
};
// Decodes a non-interlaced grey PNG (colour type 0, 8 or 16 bit).  Returns
// false (with a message in *error) for anything else.
bool read_gray_png(const std::string& path, GrayImage* out, std::string* error);
// Decodes a non-interlaced 8-bit RGB PNG (colour type 2) into rgb bytes.
bool read_rgb_png(const std::string& path, int* width, int* height, std::vector<uint8_t>* rgb,
                  std::string* error);

// Image files of a directory (extensions .jpg .jpeg .png .tiff), sorted by
// canonical path (utils.cpp:122-145).
std::set<std::string> files_directory(const std::string& data_path);

// One RGB-D pair per call: depth in metres (raw * depth_scale) in device
// memory, pitched like a cv::cuda::GpuMat of CV_32F.
class RGBDDataset {
public:
    RGBDDataset(const std::string& data_path, double depth_scale);
    ~RGBDDataset();
    // false when the files are exhausted or the depth file cannot be read;
    // with `color`, the colour file (an 8-bit RGB PNG of the depth's size)
    // is decoded too, into device memory (3 bytes per pixel)
    bool get_next(PtrStepSz<float>* depth, PtrStepSz<unsigned char>* color = nullptr);
    int num_files() const { return (int)files.size() / 2; }

private:
    std::vector<std::string> files;
    size_t next = 0;
    double depth_scale;
    float* device_depth = nullptr;
    size_t device_bytes = 0;
    unsigned char* device_color = nullptr;
    size_t device_color_bytes = 0;
};

struct Mesh {
    std::vector<float3> vertices;
    std::vector<float3> normals;  // empty, or one per vertex (welded meshes)
    std::vector<uchar3> colors;   // empty, or one per vertex (colour mode)
    std::vector<int3> polygons;
    // with normals when there are any (mesh.cpp:59-64 switches the same way)
    bool store_to_ply(const std::string& ply_filepath) const;
};

class TSDFSurface {
public:
    explicit TSDFSurface(const Config& cfg);
    TSDFVolumeConfig build_cfg() const;
    // world->camera pose {qw,qx,qy,qz,tx,ty,tz}
    void set_global_transformation(const std::array<double, 7>& pose) { global_pose = pose; }
    std::array<double, 7> get_global_transformation() const { return global_pose; }
    // One frame: integrate, marching cubes; the mesh is kept.  Colour may be
    // empty.  Returns false on failure (logged).
    template <class Color, class Depth>
    bool reconstruct_surface(const Color& color, const Depth& depth) {
        return reconstruct(PtrStepSz<float>{(float*)depth.data, (size_t)depth.step, (int)depth.rows,
                                            (int)depth.cols},
                           color.data ? (const unsigned char*)color.data : nullptr, color.step);
    }
    bool reconstruct_surface(const PtrStepSz<float>& depth) { return reconstruct(depth, nullptr, 0); }
    // tsdf_surface_reconstruction.cpp:94-120: renders the cloud at the global
    // pose into a camera_width x camera_height depth and colour pair
    // (cudaPointCloud::project_to_color_depth), integrates it and runs
    // marching cubes.  The reference stops after the render (its integrate
    // and meshing calls are commented out, :114-118) and computes normals
    // that nothing reads.  A null cloud fails (logged).
    bool reconstruct_surface(const cudaPointCloud::Ptr cuda_pc);
    bool reconstruct_surface(cudaPointCloud& cloud);
    const Mesh& get_mesh() const { return surface_mesh; }

private:
    bool reconstruct(const PtrStepSz<float>& depth, const unsigned char* rgb, size_t rgb_step);
    TSDFVolumeConfig cfg_;
    bool welded_ = false;  // parameter tsdf.mesh.welded (this library's addition)
    bool color_ = false;   // parameter dataset.color.mode: integrate colour, colour the mesh
    std::array<double, 7> global_pose{1, 0, 0, 0, 0, 0, 0};
    Mesh surface_mesh;
};

// World->camera quaternion pose from the parameter file's Euler angles in
// degrees and translation: R = Rz(roll) Ry(yaw) Rx(pitch)
// (tsdf_surface_reconstruction.cpp:66-79).
std::array<double, 7> euler_pose(double roll_deg, double yaw_deg, double pitch_deg, double tx,
                                 double ty, double tz);

}  // namespace cuphoto

// C entry points of the driver pieces (for bindings that cannot call C++).
extern "C" {
/* Decodes a grey PNG (8 or 16 bit) into pixels (capacity samples; may be
 * NULL to query the size). */
RECON_API int recon_png_read_gray(const char* path, int32_t* width, int32_t* height,
                                  int32_t* bit_depth, uint16_t* pixels, int64_t capacity);
/* Decodes an 8-bit RGB PNG into pixels (3 bytes each; capacity bytes; may be
 * NULL to query the size). */
RECON_API int recon_png_read_rgb(const char* path, int32_t* width, int32_t* height,
                                 uint8_t* pixels, int64_t capacity);
/* The value of `key` in parameter-file text. */
RECON_API int recon_config_lookup(const char* text, const char* key, char* buf, int64_t capacity);
/* euler_pose() into pose[7]. */
RECON_API void recon_euler_pose(double roll_deg, double yaw_deg, double pitch_deg, double tx,
                                double ty, double tz, double pose[7]);
}

