#include "hdr.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iterator>

namespace mpt {

namespace {

void decode(const uint8_t* p, float* out) {
    if (p[3] == 0) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    const float f = std::ldexp(1.0f, static_cast<int>(p[3]) - 136);
    out[0] = static_cast<float>(p[0]) * f;
    out[1] = static_cast<float>(p[1]) * f;
    out[2] = static_cast<float>(p[2]) * f;
}

}  // namespace

bool read_hdr(const std::string& path, std::vector<float>* rgb, int* width, int* height, std::string* err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        *err = "cannot open " + path;
        return false;
    }
    const std::vector<uint8_t> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t pos = 0;
    auto line = [&]() {
        std::string s;
        while (pos < data.size() && data[pos] != '\n') s.push_back(static_cast<char>(data[pos++]));
        if (pos < data.size()) ++pos;
        return s;
    };
    const std::string magic = line();
    if (magic.rfind("#?", 0) != 0) {
        *err = path + ": not a Radiance file";
        return false;
    }
    bool rgbe = false;
    while (pos < data.size()) {
        const std::string s = line();
        if (s.empty()) break;
        if (s == "FORMAT=32-bit_rle_rgbe") rgbe = true;
    }
    if (!rgbe) {
        *err = path + ": unsupported format (32-bit_rle_rgbe expected)";
        return false;
    }
    int w = 0, h = 0;
    if (std::sscanf(line().c_str(), "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0) {
        *err = path + ": unsupported resolution line (-Y H +X W expected)";
        return false;
    }
    // the smallest encoding of a scanline: 4 bytes a pixel flat, or (widths 8 to
    // 32767) a 4-byte header and per channel runs of up to 127 pixels in 2 bytes;
    // a header whose scanlines cannot fit in the rest of the file is rejected
    // before anything is allocated
    const size_t wz = static_cast<size_t>(w);
    const size_t min_line = (w >= 8 && w < 32768) ? std::min(4 * wz, 4 + 8 * ((wz + 126) / 127)) : 4 * wz;
    if (static_cast<uint64_t>(h) * min_line > data.size() - pos || wz * static_cast<size_t>(h) > (size_t{1} << 30)) {
        *err = path + ": resolution larger than the file holds";
        return false;
    }
    rgb->assign(static_cast<size_t>(w) * h * 3, 0.0f);
    std::vector<uint8_t> scan(static_cast<size_t>(w) * 4);
    auto need = [&](size_t n) { return pos + n <= data.size(); };
    for (int y = 0; y < h; ++y) {
        float* row = rgb->data() + static_cast<size_t>(y) * w * 3;
        const bool rle = w >= 8 && w < 32768 && need(4) && data[pos] == 2 && data[pos + 1] == 2 &&
                         (data[pos + 2] & 0x80) == 0;
        if (!rle) {
            if (!need(static_cast<size_t>(w) * 4)) {
                *err = path + ": truncated";
                return false;
            }
            for (int x = 0; x < w; ++x) decode(&data[pos + 4 * static_cast<size_t>(x)], row + 3 * x);
            pos += static_cast<size_t>(w) * 4;
            continue;
        }
        if (((data[pos + 2] << 8) | data[pos + 3]) != w) {
            *err = path + ": scanline width mismatch";
            return false;
        }
        pos += 4;
        for (int c = 0; c < 4; ++c) {
            int x = 0;
            while (x < w) {
                if (!need(1)) {
                    *err = path + ": truncated";
                    return false;
                }
                int n = data[pos++];
                if (n > 128) {
                    n -= 128;
                    if (!need(1) || x + n > w) {
                        *err = path + ": bad run";
                        return false;
                    }
                    const uint8_t v = data[pos++];
                    for (int k = 0; k < n; ++k) scan[4 * static_cast<size_t>(x++) + c] = v;
                } else {
                    if (n == 0 || !need(static_cast<size_t>(n)) || x + n > w) {
                        *err = path + ": bad run";
                        return false;
                    }
                    for (int k = 0; k < n; ++k) scan[4 * static_cast<size_t>(x++) + c] = data[pos++];
                }
            }
        }
        for (int x = 0; x < w; ++x) decode(&scan[4 * static_cast<size_t>(x)], row + 3 * x);
    }
    *width = w;
    *height = h;
    return true;
}

}  // namespace mpt
