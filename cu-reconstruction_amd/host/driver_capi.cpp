// driver_capi.cpp -- C entry points of the driver pieces (parameter file,
// depth PNG decoding, the parameter file's camera pose), for bindings that
// cannot call C++.
#include "recon/tsdf_surface.hpp"

#include <cstring>

extern "C" {

RECON_API int recon_png_read_gray(const char* path, int32_t* width, int32_t* height,
                                  int32_t* bit_depth, uint16_t* pixels, int64_t capacity) {
    if (!path || !width || !height) return RECON_ERR_NULL;
    cuphoto::GrayImage img;
    std::string err;
    if (!cuphoto::read_gray_png(path, &img, &err)) return RECON_ERR_CONFIG;
    *width = img.width;
    *height = img.height;
    if (bit_depth) *bit_depth = img.bit_depth;
    if (!pixels) return RECON_OK;
    if ((int64_t)img.pixels.size() > capacity) return RECON_ERR_CAPACITY;
    std::memcpy(pixels, img.pixels.data(), img.pixels.size() * sizeof(uint16_t));
    return RECON_OK;
}

RECON_API int recon_png_read_rgb(const char* path, int32_t* width, int32_t* height,
                                 uint8_t* pixels, int64_t capacity) {
    if (!path || !width || !height) return RECON_ERR_NULL;
    int w = 0, h = 0;
    std::vector<uint8_t> rgb;
    std::string err;
    if (!cuphoto::read_rgb_png(path, &w, &h, &rgb, &err)) return RECON_ERR_CONFIG;
    *width = w;
    *height = h;
    if (!pixels) return RECON_OK;
    if ((int64_t)rgb.size() > capacity) return RECON_ERR_CAPACITY;
    std::memcpy(pixels, rgb.data(), rgb.size());
    return RECON_OK;
}

// Value of `key` in parameter-file text (into buf, NUL-terminated);
// RECON_ERR_RANGE when absent.
RECON_API int recon_config_lookup(const char* text, const char* key, char* buf, int64_t capacity) {
    if (!text || !key || !buf || capacity <= 0) return RECON_ERR_NULL;
    cuphoto::Config cfg;
    if (!cfg.parse(text)) return RECON_ERR_CONFIG;
    if (!cfg.has(key)) return RECON_ERR_RANGE;
    const std::string v = cfg.get<std::string>(key);
    if ((int64_t)v.size() + 1 > capacity) return RECON_ERR_CAPACITY;
    std::memcpy(buf, v.c_str(), v.size() + 1);
    return RECON_OK;
}

RECON_API void recon_euler_pose(double roll_deg, double yaw_deg, double pitch_deg, double tx,
                                double ty, double tz, double pose[7]) {
    const auto p = cuphoto::euler_pose(roll_deg, yaw_deg, pitch_deg, tx, ty, tz);
    for (int i = 0; i < 7; ++i) pose[i] = p[i];
}

}  // extern "C"
