// Image files for the two pixel formats the path produces (host code):
// 8-bit RGBA as PNG (zlib-compressed IDAT) and 32-bit float RGBA as a
// single-part scanline OpenEXR file without compression.  This stands in for
// the reference's SaveImage (ImageIO.cpp:119-160, OpenImageIO), with its
// argument checks: an empty path, a path without extension and a zero-sized
// image are errors.  Both writers take the frame buffers' row order (row 0 =
// bottom) when asked and write the top row first, as the reference's PNG save
// does (Renderer.cpp:649-675).
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include <zlib.h>

#include "image_io.hpp"

namespace pt {
namespace {

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}
template <class T>
void put_le(std::vector<uint8_t>& v, T x) {
    uint8_t b[sizeof(T)];
    std::memcpy(b, &x, sizeof(T));
    v.insert(v.end(), b, b + sizeof(T));
}
void put_str(std::vector<uint8_t>& v, const char* s) { v.insert(v.end(), s, s + std::strlen(s) + 1); }

void png_chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* data, size_t n) {
    put_be32(out, (uint32_t)n);
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + n);
    put_be32(out, (uint32_t)crc32(0L, out.data() + at, (uInt)(n + 4)));
}

ImageStatus write_file(const std::string& path, const std::vector<uint8_t>& bytes, std::string& error) {
    std::error_code ec;
    const std::filesystem::path parent = std::filesystem::path(path).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent, ec);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        error = "cannot open " + path + " for writing";
        return ImageStatus::Io;
    }
    const size_t wrote = std::fwrite(bytes.data(), 1, bytes.size(), f);
    const bool closed = std::fclose(f) == 0;
    if (wrote != bytes.size() || !closed) {
        error = "short write to " + path;
        return ImageStatus::Io;
    }
    return ImageStatus::Ok;
}

}  // namespace

ImageStatus write_png_rgba8(const std::string& path, uint32_t w, uint32_t h, const uint32_t* px, bool bottom_up,
                            std::string& error) {
    // scanlines with filter byte 0, top row first
    std::vector<uint8_t> raw((size_t)h * (1 + 4 * (size_t)w));
    for (uint32_t row = 0; row < h; row++) {
        const uint32_t y = bottom_up ? h - 1 - row : row;
        uint8_t* o = raw.data() + (size_t)row * (1 + 4 * (size_t)w);
        o[0] = 0;
        std::memcpy(o + 1, px + (size_t)y * w, 4 * (size_t)w);   // bytes R, G, B, A of little-endian RGBA8 words
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 3) != Z_OK) {
        error = "PNG compression failed";
        return ImageStatus::Io;
    }
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});     // 8 bits, RGBA, deflate, adaptive filters, no interlace
    png_chunk(png, "IHDR", ihdr.data(), ihdr.size());
    png_chunk(png, "IDAT", z.data(), zn);
    png_chunk(png, "IEND", nullptr, 0);
    return write_file(path, png, error);
}

ImageStatus write_exr_rgba32f(const std::string& path, uint32_t w, uint32_t h, const float* px, bool bottom_up,
                              std::string& error) {
    std::vector<uint8_t> f = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};   // magic, version 2, single-part scanline
    auto attr = [&](const char* name, const char* type, const std::vector<uint8_t>& value) {
        put_str(f, name);
        put_str(f, type);
        put_le<int32_t>(f, (int32_t)value.size());
        f.insert(f.end(), value.begin(), value.end());
    };
    std::vector<uint8_t> v;
    for (const char* c : {"A", "B", "G", "R"}) {   // channel list in name order
        put_str(v, c);
        put_le<int32_t>(v, 2);                  // FLOAT
        v.insert(v.end(), {0, 0, 0, 0});        // pLinear, reserved
        put_le<int32_t>(v, 1);                  // x sampling
        put_le<int32_t>(v, 1);                  // y sampling
    }
    v.push_back(0);
    attr("channels", "chlist", v);
    attr("compression", "compression", {0});    // NO_COMPRESSION
    v.clear();
    for (int32_t x : {0, 0, (int32_t)w - 1, (int32_t)h - 1}) put_le<int32_t>(v, x);
    attr("dataWindow", "box2i", v);
    attr("displayWindow", "box2i", v);
    attr("lineOrder", "lineOrder", {0});        // INCREASING_Y
    v.clear();
    put_le<float>(v, 1.0f);
    attr("pixelAspectRatio", "float", v);
    v.clear();
    put_le<float>(v, 0.0f);
    put_le<float>(v, 0.0f);
    attr("screenWindowCenter", "v2f", v);
    v.clear();
    put_le<float>(v, 1.0f);
    attr("screenWindowWidth", "float", v);
    f.push_back(0);                              // end of header

    // offset table (one scanline per chunk), then the chunks: y, byte count,
    // the row's A, B, G and R values
    const size_t line_bytes = 16 * (size_t)w;
    const size_t table = f.size();
    const size_t first = table + 8 * (size_t)h;
    f.reserve(first + (size_t)h * (8 + line_bytes));
    for (uint32_t y = 0; y < h; y++) put_le<uint64_t>(f, (uint64_t)(first + (size_t)y * (8 + line_bytes)));
    static const int kChannel[4] = {3, 2, 1, 0};   // A, B, G, R from RGBA
    for (uint32_t y = 0; y < h; y++) {
        const uint32_t src = bottom_up ? h - 1 - y : y;
        put_le<int32_t>(f, (int32_t)y);
        put_le<int32_t>(f, (int32_t)line_bytes);
        const float* row = px + 4 * (size_t)src * w;
        for (int c : kChannel)
            for (uint32_t x = 0; x < w; x++) put_le<float>(f, row[4 * (size_t)x + c]);
    }
    return write_file(path, f, error);
}

ImageStatus image_kind(const std::string& path, uint32_t w, uint32_t h, ImageKind& kind, std::string& error) {
    if (path.empty()) {
        error = "Failed to save image, file path is empty";
        return ImageStatus::InvalidArgument;
    }
    std::string ext = std::filesystem::path(path).extension().string();
    if (ext.empty()) {
        error = "Failed to save image, file extension is missing, file path: " + path;
        return ImageStatus::InvalidArgument;
    }
    if (w == 0 || h == 0) {
        error = "Failed to save image, invalid dimensions " + std::to_string(w) + "x" + std::to_string(h) +
                ", file path: " + path;
        return ImageStatus::InvalidArgument;
    }
    std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    if (ext == ".png") {
        kind = ImageKind::Png;
    } else if (ext == ".exr") {
        kind = ImageKind::Exr;
    } else {
        error = "no image writer for the extension " + ext + " (png and exr are written), file path: " + path;
        return ImageStatus::Unsupported;
    }
    return ImageStatus::Ok;
}

}  // namespace pt
