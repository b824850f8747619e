// png_reader.cpp -- decoding of the depth images of an RGB-D directory
// (the reference reads them with cv::imread(..., IMREAD_ANYDEPTH),
// examples/common/rgbd_dataset.cpp:26) and, in colour mode, of the colour
// images.  Grey or RGB PNGs of 8 or 16 bits, not interlaced: chunks, zlib
// inflate of the IDAT stream (zlib), and the five scanline filters of the
// PNG specification.
#include "recon/tsdf_surface.hpp"

#include <zlib.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>

namespace cuphoto {

namespace {

uint32_t be32(const unsigned char* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

bool fail(std::string* error, const std::string& what) {
    if (error) *error = what;
    return false;
}

}  // namespace

static bool read_png(const std::string& path, GrayImage* out, std::string* error);

bool read_gray_png(const std::string& path, GrayImage* out, std::string* error) {
    if (!read_png(path, out, error)) return false;
    if (out->channels != 1) return fail(error, path + ": not a grey PNG");
    return true;
}

bool read_rgb_png(const std::string& path, int* width, int* height, std::vector<uint8_t>* rgb,
                  std::string* error) {
    GrayImage img;
    if (!read_png(path, &img, error)) return false;
    if (img.channels != 3 || img.bit_depth != 8) return fail(error, path + ": not an 8-bit RGB PNG");
    *width = img.width;
    *height = img.height;
    rgb->resize(img.pixels.size());
    for (size_t i = 0; i < img.pixels.size(); ++i) (*rgb)[i] = (uint8_t)img.pixels[i];
    return true;
}

static bool read_png(const std::string& path, GrayImage* out, std::string* error) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return fail(error, "cannot open " + path);
    const std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (buf.size() < 8 || std::memcmp(buf.data(), sig, 8) != 0) return fail(error, path + ": not a PNG");
    size_t pos = 8;
    int width = 0, height = 0, depth = 0, ctype = -1, interlace = 0;
    std::vector<unsigned char> idat;
    bool have_header = false, ended = false;
    while (pos + 12 <= buf.size() && !ended) {
        const uint32_t len = be32(&buf[pos]);
        if (len > buf.size() - pos - 12) return fail(error, path + ": truncated chunk");
        const unsigned char* type = &buf[pos + 4];
        const unsigned char* data = &buf[pos + 8];
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (len < 13) return fail(error, path + ": bad IHDR");
            width = (int)be32(data);
            height = (int)be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
            have_header = true;
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), data, data + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            ended = true;
        }
        pos += 12 + len;
    }
    if (!have_header) return fail(error, path + ": no IHDR");
    if ((ctype != 0 && ctype != 2) || (depth != 8 && depth != 16) || interlace != 0)
        return fail(error, path + ": only non-interlaced grey or RGB PNGs of 8 or 16 bits are read");
    if (width <= 0 || height <= 0 || (int64_t)width * height > (int64_t(1) << 28))
        return fail(error, path + ": bad size");
    const int channels = ctype == 2 ? 3 : 1;
    const size_t bpp = (size_t)channels * (depth / 8);  // bytes per pixel
    const size_t stride = (size_t)width * bpp;
    std::vector<unsigned char> raw((stride + 1) * (size_t)height);
    uLongf raw_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &raw_len, idat.data(), (uLong)idat.size()) != Z_OK ||
        raw_len != raw.size())
        return fail(error, path + ": corrupt image data");
    std::vector<unsigned char> prev(stride, 0), cur(stride);
    out->width = width;
    out->height = height;
    out->bit_depth = depth;
    out->channels = channels;
    const size_t samples = (size_t)width * channels;  // per row
    out->pixels.assign(samples * height, 0);
    for (int y = 0; y < height; ++y) {
        const unsigned char* line = &raw[(size_t)y * (stride + 1)];
        const int filter = line[0];
        for (size_t i = 0; i < stride; ++i) {
            const int x = line[1 + i];
            const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v;
            switch (filter) {
                case 0: v = x; break;
                case 1: v = x + a; break;
                case 2: v = x + b; break;
                case 3: v = x + ((a + b) >> 1); break;
                case 4: v = x + paeth(a, b, c); break;
                default: return fail(error, path + ": bad filter type");
            }
            cur[i] = (unsigned char)v;
        }
        uint16_t* row = &out->pixels[(size_t)y * samples];
        for (size_t x = 0; x < samples; ++x)
            row[x] = depth == 16 ? (uint16_t)((cur[2 * x] << 8) | cur[2 * x + 1]) : cur[x];
        prev.swap(cur);
    }
    return true;
}

}  // namespace cuphoto
