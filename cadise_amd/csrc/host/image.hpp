// Picture files for image textures: PNG and TGA (8-bit data) and Radiance
// .hdr (float data), decoded here and turned into texture texels the way the
// reference's PictureLoader::loadRgbImage (pictureLoader.cpp:14-176) does.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mpt {

// A decoded 8-bit picture, top row first: its colour samples (1 channel
// for gray, 3 for RGB and palette pictures; 16-bit samples keep their high
// byte, 1/2/4-bit gray is scaled to 0..255) and whether the file carries
// alpha (an alpha channel or a tRNS chunk; the alpha samples themselves are
// not kept).
struct Picture8 {
    int32_t width = 0, height = 0, channels = 0;
    bool has_alpha = false;
    std::vector<uint8_t> data;   // width * height * channels
};

// PNG (all colour types and bit depths, Adam7 interlacing) and TGA (types 2,
// 3, 10, 11 at 8, 24 or 32 bits per pixel).
bool decode_png(const std::vector<uint8_t>& file, Picture8* out, std::string* err);
bool decode_tga(const std::vector<uint8_t>& file, Picture8* out, std::string* err);

// PictureLoader::loadRgbImage: the texels (3 channels, row 0 = the picture's
// bottom row, as stbi_set_flip_vertically_on_load leaves them) of a .png /
// .tga (8-bit: each sample v becomes k / 255 with k = the byte of
// inverse_gamma_correction(v / 255) * 255 + 0.5; gray repeated over the three
// channels; a picture with alpha gives all zeros) or .hdr file (floats as
// read). Other extensions (.jpg among them) are rejected.
bool load_rgb_image(const std::string& path, int32_t* width, int32_t* height, std::vector<float>* rgb,
                    std::string* err);

// Whether load_rgb_image reads files of this name's extension.
bool rgb_image_supported(const std::string& path);

}  // namespace mpt
