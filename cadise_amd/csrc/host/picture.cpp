// 8-bit picture writers of the film save: PPM, PNG and baseline JPEG.
//
// PictureSaver::save (pictureSaver.cpp:12-82) picks the writer by the file
// extension: PPM through its own fprintf / fwrite, PNG and JPG (quality 92)
// through the vendored stb_image_write. These writers are this library's own:
//   PPM   the reference's header "P6 <w> <h> 255\n" and the raw RGB payload, so
//         the file is byte-identical to savePPM's;
//   PNG   8-bit RGB, filter type 0 per row, zlib stream of stored deflate
//         blocks (RFC 1950 / 1951) and CRC-32 chunks (PNG specification);
//   JPEG  baseline sequential DCT (ITU-T T.81): JFIF YCbCr (CCIR 601
//         weights), no chroma subsampling, the DCT-II of T.81 A.3.3 evaluated
//         from its definition in double precision, the quantisation tables of
//         T.81 Annex K.1 scaled for quality 92 by the IJG quality rule, and the
//         Huffman tables of Annex K.3.
// Rows are written top to bottom, as the film holds them.
#include "picture.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

namespace mpt {
namespace {

bool write_file(const std::string& path, const std::vector<uint8_t>& bytes, std::string* err) {
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        *err = "cannot write " + path;
        return false;
    }
    f.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
    f.close();
    if (!f) {
        *err = "short write " + path;
        return false;
    }
    return true;
}

bool has_ext(const std::string& p, const char* e) {
    const size_t n = std::strlen(e);
    return p.size() >= n && p.compare(p.size() - n, n, e) == 0;
}

// ---- PPM ----------------------------------------------------------------------------

std::vector<uint8_t> encode_ppm(const uint8_t* rgb, int w, int h) {
    char head[64];
    const int n = std::snprintf(head, sizeof(head), "P6 %d %d 255\n", w, h);
    std::vector<uint8_t> out(head, head + n);
    out.insert(out.end(), rgb, rgb + static_cast<size_t>(3) * w * h);
    return out;
}

// ---- PNG ----------------------------------------------------------------------------

uint32_t crc32_of(const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t k = i;
            for (int j = 0; j < 8; ++j) k = (k & 1) ? 0xEDB88320u ^ (k >> 1) : k >> 1;
            table[i] = k;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back(static_cast<uint8_t>(x >> 24));
    v.push_back(static_cast<uint8_t>(x >> 16));
    v.push_back(static_cast<uint8_t>(x >> 8));
    v.push_back(static_cast<uint8_t>(x));
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put_be32(out, static_cast<uint32_t>(data.size()));
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc32_of(out.data() + at, out.size() - at));
}

std::vector<uint8_t> encode_png(const uint8_t* rgb, int w, int h) {
    const size_t row = static_cast<size_t>(3) * w;
    std::vector<uint8_t> raw;
    raw.reserve(static_cast<size_t>(h) * (1 + row));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb + y * row, rgb + (y + 1) * row);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521;
        b = (b + a) % 65521;
    }
    size_t pos = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(static_cast<uint8_t>(n));
        z.push_back(static_cast<uint8_t>(n >> 8));
        z.push_back(static_cast<uint8_t>(~n));
        z.push_back(static_cast<uint8_t>(~n >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
        pos += n;
    } while (pos < raw.size());
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, static_cast<uint32_t>(w));
    put_be32(ihdr, static_cast<uint32_t>(h));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    png_chunk(png, "IHDR", ihdr);
    png_chunk(png, "IDAT", z);
    png_chunk(png, "IEND", {});
    return png;
}

// ---- baseline JPEG ------------------------------------------------------------------

// T.81 Annex K.1, Tables K.1 (luminance) and K.2 (chrominance), row-major.
const uint8_t kLumQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                           14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                           18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                           49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kChromQ[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                             24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                             99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                             99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};

// T.81 Annex K.3: code-length counts (lengths 1..16) and symbols of the typical
// DC and AC tables (Tables K.3 - K.6).
const uint8_t kDcLumBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChromBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcLumVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
const uint8_t kAcChromBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
const uint8_t kAcChromVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

// Zig-zag scan order (T.81 Figure A.6): zz[k] = row-major index of the k-th
// coefficient, by walking the anti-diagonals in alternating directions.
struct ZigZag {
    int zz[64];
    ZigZag() {
        int k = 0;
        for (int s = 0; s < 15; ++s) {
            if (s % 2 == 0) {
                for (int r = std::min(s, 7); r >= 0 && s - r <= 7; --r) zz[k++] = 8 * r + (s - r);
            } else {
                for (int c = std::min(s, 7); c >= 0 && s - c <= 7; --c) zz[k++] = 8 * (s - c) + c;
            }
        }
    }
};

// Huffman code per symbol from the code-length counts (T.81 Annex C).
struct HuffCode {
    uint16_t code[256];
    uint8_t  len[256];
    HuffCode(const uint8_t* bits, const uint8_t* vals) {
        std::memset(len, 0, sizeof(len));
        uint16_t c = 0;
        int k = 0;
        for (int l = 1; l <= 16; ++l) {
            for (int i = 0; i < bits[l - 1]; ++i, ++k) {
                code[vals[k]] = c++;
                len[vals[k]] = static_cast<uint8_t>(l);
            }
            c <<= 1;
        }
    }
};

// The IJG quality rule: scale = 5000 / q below 50, else 200 - 2q (percent);
// entries (base * scale + 50) / 100 clamped to [1, 255].
void quality_table(const uint8_t* base, int quality, uint8_t* out) {
    const int q = quality < 1 ? 1 : (quality > 100 ? 100 : quality);
    const int scale = q < 50 ? 5000 / q : 200 - 2 * q;
    for (int i = 0; i < 64; ++i) {
        int v = (base[i] * scale + 50) / 100;
        out[i] = static_cast<uint8_t>(v < 1 ? 1 : (v > 255 ? 255 : v));
    }
}

struct BitWriter {
    std::vector<uint8_t>& out;
    uint32_t acc = 0;
    int      n = 0;
    explicit BitWriter(std::vector<uint8_t>& o) : out(o) {}
    void put(uint32_t bits, int len) {
        for (int i = len - 1; i >= 0; --i) {
            acc = (acc << 1) | ((bits >> i) & 1u);
            if (++n == 8) {
                out.push_back(static_cast<uint8_t>(acc));
                if (acc == 0xFF) out.push_back(0x00);   // byte stuffing (T.81 F.1.2.3)
                acc = 0;
                n = 0;
            }
        }
    }
    void flush() {   // pad with 1-bits (T.81 F.1.2.3)
        while (n != 0) put(1, 1);
    }
};

// Magnitude category and the additional bits of a coefficient (T.81 F.1.2.1).
inline void category(int v, int* size, uint32_t* bits) {
    int a = v < 0 ? -v : v;
    int s = 0;
    while (a) {
        ++s;
        a >>= 1;
    }
    *size = s;
    *bits = static_cast<uint32_t>(v < 0 ? v + (1 << s) - 1 : v) & ((1u << s) - 1u);
}

void put_marker(std::vector<uint8_t>& o, uint8_t m) {
    o.push_back(0xFF);
    o.push_back(m);
}
void put_be16(std::vector<uint8_t>& o, int v) {
    o.push_back(static_cast<uint8_t>(v >> 8));
    o.push_back(static_cast<uint8_t>(v));
}

void put_dht(std::vector<uint8_t>& o, int cls_id, const uint8_t* bits, const uint8_t* vals) {
    int n = 0;
    for (int i = 0; i < 16; ++i) n += bits[i];
    put_marker(o, 0xC4);
    put_be16(o, 2 + 1 + 16 + n);
    o.push_back(static_cast<uint8_t>(cls_id));
    o.insert(o.end(), bits, bits + 16);
    o.insert(o.end(), vals, vals + n);
}

std::vector<uint8_t> encode_jpeg(const uint8_t* rgb, int w, int h, int quality) {
    static const ZigZag Z;
    static const HuffCode dc_l(kDcLumBits, kDcVals), dc_c(kDcChromBits, kDcVals);
    static const HuffCode ac_l(kAcLumBits, kAcLumVals), ac_c(kAcChromBits, kAcChromVals);
    // cos((2x + 1) u pi / 16) * C(u) / 2, so that F = sum_x sum_y M[u][x] M[v][y] f
    static double M[8][8];
    static bool init = false;
    if (!init) {
        for (int u = 0; u < 8; ++u)
            for (int x = 0; x < 8; ++x)
                M[u][x] = (u == 0 ? std::sqrt(0.5) : 1.0) * 0.5 * std::cos((2 * x + 1) * u * M_PI / 16.0);
        init = true;
    }
    uint8_t q[2][64];
    quality_table(kLumQ, quality, q[0]);
    quality_table(kChromQ, quality, q[1]);

    std::vector<uint8_t> o;
    put_marker(o, 0xD8);   // SOI
    put_marker(o, 0xE0);   // APP0 JFIF 1.01, no density units, 1:1
    put_be16(o, 16);
    o.insert(o.end(), {'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0});
    for (int t = 0; t < 2; ++t) {   // DQT, 8-bit entries in zig-zag order
        put_marker(o, 0xDB);
        put_be16(o, 2 + 1 + 64);
        o.push_back(static_cast<uint8_t>(t));
        for (int k = 0; k < 64; ++k) o.push_back(q[t][Z.zz[k]]);
    }
    put_marker(o, 0xC0);   // SOF0: 8-bit precision, 3 components, 1 x 1 sampling each
    put_be16(o, 8 + 3 * 3);
    o.push_back(8);
    put_be16(o, h);
    put_be16(o, w);
    o.push_back(3);
    for (int c = 0; c < 3; ++c) {
        o.push_back(static_cast<uint8_t>(c + 1));
        o.push_back(0x11);
        o.push_back(static_cast<uint8_t>(c == 0 ? 0 : 1));
    }
    put_dht(o, 0x00, kDcLumBits, kDcVals);
    put_dht(o, 0x10, kAcLumBits, kAcLumVals);
    put_dht(o, 0x01, kDcChromBits, kDcVals);
    put_dht(o, 0x11, kAcChromBits, kAcChromVals);
    put_marker(o, 0xDA);   // SOS: all three components interleaved, full spectral range
    put_be16(o, 6 + 2 * 3);
    o.push_back(3);
    for (int c = 0; c < 3; ++c) {
        o.push_back(static_cast<uint8_t>(c + 1));
        o.push_back(static_cast<uint8_t>(c == 0 ? 0x00 : 0x11));
    }
    o.insert(o.end(), {0, 63, 0});

    BitWriter bw(o);
    int pred[3] = {0, 0, 0};
    double blk[3][64], tmp[64];
    for (int by = 0; by < h; by += 8) {
        for (int bx = 0; bx < w; bx += 8) {
            // colour conversion and level shift; edge blocks repeat the last row / column
            for (int y = 0; y < 8; ++y) {
                const int sy = std::min(by + y, h - 1);
                for (int x = 0; x < 8; ++x) {
                    const int sx = std::min(bx + x, w - 1);
                    const uint8_t* p = rgb + 3 * (static_cast<size_t>(sy) * w + sx);
                    const double r = p[0], g = p[1], b = p[2];
                    blk[0][8 * y + x] = 0.299 * r + 0.587 * g + 0.114 * b - 128.0;
                    blk[1][8 * y + x] = -0.168736 * r - 0.331264 * g + 0.5 * b;
                    blk[2][8 * y + x] = 0.5 * r - 0.418688 * g - 0.081312 * b;
                }
            }
            for (int c = 0; c < 3; ++c) {
                // separable DCT-II: rows, then columns
                for (int y = 0; y < 8; ++y)
                    for (int u = 0; u < 8; ++u) {
                        double s = 0.0;
                        for (int x = 0; x < 8; ++x) s += M[u][x] * blk[c][8 * y + x];
                        tmp[8 * y + u] = s;
                    }
                int coef[64];
                for (int v = 0; v < 8; ++v)
                    for (int u = 0; u < 8; ++u) {
                        double s = 0.0;
                        for (int y = 0; y < 8; ++y) s += M[v][y] * tmp[8 * y + u];
                        coef[8 * v + u] = static_cast<int>(std::lround(s / q[c == 0 ? 0 : 1][8 * v + u]));
                    }
                const HuffCode& dc = c == 0 ? dc_l : dc_c;
                const HuffCode& ac = c == 0 ? ac_l : ac_c;
                int size;
                uint32_t bits;
                const int diff = coef[0] - pred[c];
                pred[c] = coef[0];
                category(diff, &size, &bits);
                bw.put(dc.code[size], dc.len[size]);
                if (size) bw.put(bits, size);
                int run = 0;
                for (int k = 1; k < 64; ++k) {
                    const int v = coef[Z.zz[k]];
                    if (v == 0) {
                        ++run;
                        continue;
                    }
                    while (run > 15) {
                        bw.put(ac.code[0xF0], ac.len[0xF0]);   // ZRL
                        run -= 16;
                    }
                    category(v, &size, &bits);
                    const int sym = (run << 4) | size;
                    bw.put(ac.code[sym], ac.len[sym]);
                    bw.put(bits, size);
                    run = 0;
                }
                if (run) bw.put(ac.code[0x00], ac.len[0x00]);   // EOB
            }
        }
    }
    bw.flush();
    put_marker(o, 0xD9);   // EOI
    return o;
}

}  // namespace

int picture_kind(const std::string& path) {
    if (has_ext(path, ".ppm") || has_ext(path, ".PPM")) return PICTURE_PPM;
    if (has_ext(path, ".png") || has_ext(path, ".PNG")) return PICTURE_PNG;
    if (has_ext(path, ".jpg") || has_ext(path, ".JPG") || has_ext(path, ".jpeg") || has_ext(path, ".JPEG"))
        return PICTURE_JPEG;
    if (has_ext(path, ".pfm") || has_ext(path, ".PFM")) return PICTURE_PFM;
    return PICTURE_NONE;
}

bool write_picture8(const std::string& path, const uint8_t* rgb, int w, int h, std::string* err) {
    switch (picture_kind(path)) {
        case PICTURE_PPM: return write_file(path, encode_ppm(rgb, w, h), err);
        case PICTURE_PNG: return write_file(path, encode_png(rgb, w, h), err);
        case PICTURE_JPEG: return write_file(path, encode_jpeg(rgb, w, h, kJpegQuality), err);
        default: *err = "output format of " + path + " (png, jpg, ppm; pfm for linear films)"; return false;
    }
}

bool write_pfm(const std::string& path, const float* film, int w, int h, std::string* err) {
    std::vector<uint8_t> out;
    char head[64];
    const int n = std::snprintf(head, sizeof(head), "PF\n%d %d\n-1.0\n", w, h);
    out.insert(out.end(), head, head + n);
    out.reserve(out.size() + static_cast<size_t>(12) * w * h);
    // PFM rows run bottom to top; little-endian floats (scale -1)
    for (int y = h - 1; y >= 0; --y)
        for (int x = 0; x < w; ++x) {
            const float* px = film + 4 * (static_cast<size_t>(y) * w + x);
            const float inv = px[3] == 0.f ? 0.f : 1.f / px[3];
            const float rgb[3] = {px[0] * inv, px[1] * inv, px[2] * inv};
            const uint8_t* b = reinterpret_cast<const uint8_t*>(rgb);
            out.insert(out.end(), b, b + sizeof(rgb));
        }
    return write_file(path, out, err);
}

}  // namespace mpt

// ---- C ABI (host only: no device needed) --------------------------------------------

#include "../../../include/mpt.h"

extern "C" int mpt_image_write(const char* path, const uint8_t* rgb, int32_t width, int32_t height) {
    if (!path || width <= 0 || height <= 0 || !rgb) return MPT_ERR_INVALID;
    if (static_cast<int64_t>(width) * height > (static_cast<int64_t>(1) << 31)) return MPT_ERR_INVALID;
    const std::string p(path);
    const int kind = mpt::picture_kind(p);
    if (kind == mpt::PICTURE_NONE || kind == mpt::PICTURE_PFM) return MPT_ERR_UNSUPPORTED;
    std::string err;
    return mpt::write_picture8(p, rgb, width, height, &err) ? MPT_OK : MPT_ERR_IO;
}
