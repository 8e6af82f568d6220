// 8-bit picture writers of the film save (PPM, PNG, baseline JPEG) and the
// linear PFM writer.
#pragma once

#include <cstdint>
#include <string>

namespace mpt {

enum PictureKind { PICTURE_NONE = 0, PICTURE_PPM = 1, PICTURE_PNG = 2, PICTURE_JPEG = 3, PICTURE_PFM = 4 };

// JPEG quality of PictureSaver::saveJPG (pictureSaver.cpp:63-72).
constexpr int kJpegQuality = 92;

// The writer a file name selects by its extension.
int picture_kind(const std::string& path);
// rgb: h rows of w x 3 bytes, top row first. False (and *err) on failure.
bool write_picture8(const std::string& path, const uint8_t* rgb, int w, int h, std::string* err);
// film: h x w x 4 floats (sum w*rgb, sum w); writes sum w*rgb / sum w.
bool write_pfm(const std::string& path, const float* film, int w, int h, std::string* err);

}  // namespace mpt
