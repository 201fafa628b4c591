// PNG and OpenEXR writers for the frame buffers (host code).
#pragma once
#include <cstdint>
#include <string>

namespace pt {

enum class ImageStatus { Ok, InvalidArgument, Unsupported, Io };
enum class ImageKind { Png, Exr };

// The container a path asks for: ".png" or ".exr" (any case); checks the
// path and size first (empty path, missing extension, zero size).
ImageStatus image_kind(const std::string& path, uint32_t width, uint32_t height, ImageKind& kind, std::string& error);
// 8-bit RGBA PNG of little-endian RGBA8 words.
ImageStatus write_png_rgba8(const std::string& path, uint32_t width, uint32_t height, const uint32_t* pixels,
                            bool bottom_up, std::string& error);
// OpenEXR scanline file, FLOAT channels A, B, G, R from float RGBA, no compression.
ImageStatus write_exr_rgba32f(const std::string& path, uint32_t width, uint32_t height, const float* pixels,
                              bool bottom_up, std::string& error);

}  // namespace pt
