#include "image.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>

#include "hdr.hpp"

namespace mpt {

namespace {

// ---- DEFLATE (RFC 1951) inside a zlib stream (RFC 1950) ---------------------------

struct Bits {
    const uint8_t* p;
    size_t n;
    size_t pos = 0;
    uint32_t buf = 0;
    int cnt = 0;
    bool bad = false;
    // next `need` bits (need <= 16), least significant first
    int get(int need) {
        while (cnt < need) {
            if (pos >= n) {
                bad = true;
                return 0;
            }
            buf |= static_cast<uint32_t>(p[pos++]) << cnt;
            cnt += 8;
        }
        const int v = static_cast<int>(buf & ((1u << need) - 1u));
        buf >>= need;
        cnt -= need;
        return v;
    }
    void align() {   // the rest of the current byte
        buf = 0;
        cnt = 0;
    }
};

// Canonical Huffman code: number of codes per length and the symbols ordered
// by (length, symbol value).
struct Huff {
    uint16_t count[16];
    uint16_t symbol[320];
};

// False when the lengths over-subscribe the code space; incomplete codes are
// accepted.
bool huff_build(Huff* h, const uint8_t* len, int n) {
    std::memset(h->count, 0, sizeof(h->count));
    for (int i = 0; i < n; ++i) h->count[len[i]]++;
    int left = 1;
    for (int l = 1; l < 16; ++l) {
        left = 2 * left - h->count[l];
        if (left < 0) return false;
    }
    uint16_t offs[16];
    offs[1] = 0;
    for (int l = 1; l < 15; ++l) offs[l + 1] = static_cast<uint16_t>(offs[l] + h->count[l]);
    for (int i = 0; i < n; ++i)
        if (len[i] != 0) h->symbol[offs[len[i]]++] = static_cast<uint16_t>(i);
    return true;
}

// One symbol: the code is read most significant bit first, one bit at a
// time, and compared with the first code of each length.
int huff_decode(Bits* b, const Huff& h) {
    int code = 0, first = 0, index = 0;
    for (int l = 1; l < 16; ++l) {
        code |= b->get(1);
        const int c = h.count[l];
        if (code - c < first) return h.symbol[index + (code - first)];
        index += c;
        first = (first + c) << 1;
        code <<= 1;
    }
    return -1;
}

const uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                               31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

bool inflate_codes(Bits* b, const Huff& lit, const Huff& dist, std::vector<uint8_t>* out) {
    for (;;) {
        int sym = huff_decode(b, lit);
        if (sym < 0 || b->bad) return false;
        if (sym < 256) {
            out->push_back(static_cast<uint8_t>(sym));
            continue;
        }
        if (sym == 256) return true;
        sym -= 257;
        if (sym >= 29) return false;
        const int len = kLenBase[sym] + b->get(kLenExtra[sym]);
        const int ds = huff_decode(b, dist);
        if (ds < 0 || ds >= 30 || b->bad) return false;
        const size_t d = static_cast<size_t>(kDistBase[ds] + b->get(kDistExtra[ds]));
        if (b->bad || d > out->size()) return false;
        const size_t from = out->size() - d;
        for (int k = 0; k < len; ++k) out->push_back((*out)[from + static_cast<size_t>(k)]);
    }
}

bool inflate_zlib(const uint8_t* p, size_t n, std::vector<uint8_t>* out, std::string* err) {
    auto bad = [&](const char* m) {
        *err = std::string("zlib stream: ") + m;
        return false;
    };
    if (n < 2) return bad("truncated");
    const int cmf = p[0], flg = p[1];
    if ((cmf * 256 + flg) % 31 != 0) return bad("bad header");
    if ((cmf & 15) != 8) return bad("not deflate");
    if (flg & 32) return bad("preset dictionary");
    Bits b{p + 2, n - 2};
    static const Huff fixed_lit = [] {
        uint8_t l[288];
        for (int i = 0; i < 288; ++i) l[i] = i < 144 ? 8 : i < 256 ? 9 : i < 280 ? 7 : 8;
        Huff h;
        huff_build(&h, l, 288);
        return h;
    }();
    static const Huff fixed_dist = [] {
        uint8_t l[30];
        std::fill(l, l + 30, 5);
        Huff h;
        huff_build(&h, l, 30);
        return h;
    }();
    bool last = false;
    while (!last) {
        last = b.get(1) != 0;
        const int type = b.get(2);
        if (b.bad) return bad("truncated");
        if (type == 0) {
            b.align();
            if (b.pos + 4 > b.n) return bad("truncated stored block");
            const int len = b.p[b.pos] | (b.p[b.pos + 1] << 8);
            const int nlen = b.p[b.pos + 2] | (b.p[b.pos + 3] << 8);
            if (len != (~nlen & 0xFFFF)) return bad("stored block length");
            b.pos += 4;
            if (b.pos + static_cast<size_t>(len) > b.n) return bad("truncated stored block");
            out->insert(out->end(), b.p + b.pos, b.p + b.pos + len);
            b.pos += static_cast<size_t>(len);
        } else if (type == 1) {
            if (!inflate_codes(&b, fixed_lit, fixed_dist, out)) return bad("corrupt fixed block");
        } else if (type == 2) {
            const int hlit = b.get(5) + 257, hdist = b.get(5) + 1, hclen = b.get(4) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {0};
            for (int i = 0; i < hclen; ++i) cl[order[i]] = static_cast<uint8_t>(b.get(3));
            Huff lh;
            if (b.bad || hlit > 286 || hdist > 30 || !huff_build(&lh, cl, 19)) return bad("corrupt dynamic header");
            uint8_t lens[320] = {0};
            int i = 0;
            while (i < hlit + hdist) {
                const int sym = huff_decode(&b, lh);
                if (sym < 0 || b.bad) return bad("corrupt code lengths");
                if (sym < 16) {
                    lens[i++] = static_cast<uint8_t>(sym);
                    continue;
                }
                int rep = 0;
                uint8_t v = 0;
                if (sym == 16) {
                    if (i == 0) return bad("repeat without a length");
                    v = lens[i - 1];
                    rep = 3 + b.get(2);
                } else if (sym == 17) {
                    rep = 3 + b.get(3);
                } else {
                    rep = 11 + b.get(7);
                }
                if (i + rep > hlit + hdist) return bad("code lengths overflow");
                while (rep--) lens[i++] = v;
            }
            if (lens[256] == 0) return bad("no end-of-block code");
            Huff lit, dist;
            if (!huff_build(&lit, lens, hlit) || !huff_build(&dist, lens + hlit, hdist)) return bad("over-subscribed code");
            if (!inflate_codes(&b, lit, dist, out)) return bad("corrupt dynamic block");
        } else {
            return bad("reserved block type");
        }
    }
    return true;
}

uint32_t be32(const uint8_t* p) {
    return (static_cast<uint32_t>(p[0]) << 24) | (static_cast<uint32_t>(p[1]) << 16) | (static_cast<uint32_t>(p[2]) << 8) |
           static_cast<uint32_t>(p[3]);
}

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return static_cast<uint8_t>(a);
    if (pb <= pc) return static_cast<uint8_t>(b);
    return static_cast<uint8_t>(c);
}

}  // namespace

bool decode_png(const std::vector<uint8_t>& f, Picture8* out, std::string* err) {
    auto bad = [&](const std::string& m) {
        *err = "PNG: " + m;
        return false;
    };
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return bad("bad signature");
    uint32_t w = 0, h = 0;
    int depth = 0, color = -1, interlace = 0;
    bool have_hdr = false, trns = false, end = false;
    std::vector<uint8_t> pal, idat;
    size_t pos = 8;
    while (!end) {
        if (pos + 12 > f.size()) return bad("truncated chunk");
        const uint32_t len = be32(&f[pos]);
        if (len > f.size() - pos - 12) return bad("truncated chunk");
        const uint8_t* t = &f[pos + 4];
        const uint8_t* d = &f[pos + 8];
        const std::string type(reinterpret_cast<const char*>(t), 4);
        if (!have_hdr && type != "IHDR" && type != "CgBI") return bad("first chunk is not IHDR");
        if (type == "IHDR") {
            if (have_hdr || len != 13) return bad("bad IHDR");
            have_hdr = true;
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            color = d[9];
            if (d[10] != 0 || d[11] != 0) return bad("unknown compression or filter method");
            interlace = d[12];
            if (interlace > 1) return bad("unknown interlace method");
            if (w == 0 || h == 0) return bad("empty image");
            const bool ok_depth = depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
            const bool ok = ok_depth && (color == 0 || ((color == 2 || color == 4 || color == 6) && depth >= 8) ||
                                         (color == 3 && depth <= 8));
            if (!ok) return bad("unsupported colour type / bit depth");
            if ((1u << 30) / w / 4 < h) return bad("image too large");
        } else if (type == "CgBI") {
            return bad("CgBI (iPhone) images are not read");
        } else if (type == "PLTE") {
            if (len % 3 != 0 || len / 3 > 256 || len == 0) return bad("bad PLTE");
            pal.assign(d, d + len);
        } else if (type == "tRNS") {
            if (!idat.empty()) return bad("tRNS after IDAT");
            trns = true;
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            end = true;
        } else if ((t[0] & 32) == 0) {
            return bad("unknown critical chunk " + type);
        }
        pos += 12 + static_cast<size_t>(len);
    }
    if (color == 3 && pal.empty()) return bad("palette image without PLTE");
    std::vector<uint8_t> raw;
    if (!inflate_zlib(idat.data(), idat.size(), &raw, err)) return false;

    const int img_n = color == 2 ? 3 : color == 4 ? 2 : color == 6 ? 4 : 1;   // samples per pixel in the file
    const int color_n = (color == 2 || color == 3 || color == 6) ? 3 : 1;   // colour channels kept
    const int bpp_bits = img_n * depth;
    static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1}, adx[7] = {8, 8, 4, 4, 2, 2, 1},
                     ady[7] = {8, 8, 8, 4, 4, 2, 2};
    {   // the filtered rows of every pass must be in the inflated data before the image is allocated
        size_t total = 0;
        for (int pass = 0; pass < (interlace ? 7 : 1); ++pass) {
            const uint32_t x0 = interlace ? ax0[pass] : 0, y0 = interlace ? ay0[pass] : 0;
            const uint32_t dx = interlace ? adx[pass] : 1, dy = interlace ? ady[pass] : 1;
            const uint64_t pw = w > x0 ? (w - x0 + dx - 1) / dx : 0, ph = h > y0 ? (h - y0 + dy - 1) / dy : 0;
            if (pw && ph) total += static_cast<size_t>(((pw * bpp_bits + 7) / 8 + 1) * ph);
        }
        if (raw.size() < total) return bad("not enough pixel data");
    }
    out->width = static_cast<int32_t>(w);
    out->height = static_cast<int32_t>(h);
    out->channels = color_n;
    out->has_alpha = color == 4 || color == 6 || trns;
    out->data.assign(static_cast<size_t>(w) * h * color_n, 0);
    const int fbytes = std::max(1, bpp_bits / 8);   // filter distance
    const int gray_scale = (color == 0 || color == 4) && depth < 8 ? (depth == 1 ? 0xFF : depth == 2 ? 0x55 : 0x11) : 1;
    size_t off = 0;
    const int npass = interlace ? 7 : 1;
    for (int pass = 0; pass < npass; ++pass) {
        const uint32_t x0 = interlace ? ax0[pass] : 0, y0 = interlace ? ay0[pass] : 0;
        const uint32_t dx = interlace ? adx[pass] : 1, dy = interlace ? ady[pass] : 1;
        const uint32_t pw = w > x0 ? (w - x0 + dx - 1) / dx : 0, ph = h > y0 ? (h - y0 + dy - 1) / dy : 0;
        if (pw == 0 || ph == 0) continue;
        const size_t stride = (static_cast<size_t>(pw) * bpp_bits + 7) / 8;
        if (raw.size() - off < (stride + 1) * ph) return bad("not enough pixel data");
        std::vector<uint8_t> prev(stride, 0), cur(stride);
        for (uint32_t y = 0; y < ph; ++y) {
            const uint8_t ft = raw[off];
            const uint8_t* src = &raw[off + 1];
            off += stride + 1;
            for (size_t i = 0; i < stride; ++i) {
                const int a = i >= static_cast<size_t>(fbytes) ? cur[i - fbytes] : 0;
                const int b = prev[i];
                const int c = i >= static_cast<size_t>(fbytes) ? prev[i - fbytes] : 0;
                int v = src[i];
                switch (ft) {
                    case 0: break;
                    case 1: v += a; break;
                    case 2: v += b; break;
                    case 3: v += (a + b) >> 1; break;
                    case 4: v += paeth(a, b, c); break;
                    default: return bad("unknown filter type");
                }
                cur[i] = static_cast<uint8_t>(v);
            }
            for (uint32_t x = 0; x < pw; ++x) {
                uint8_t s[4];
                for (int k = 0; k < img_n; ++k) {
                    if (depth == 8) {
                        s[k] = cur[static_cast<size_t>(x) * img_n + k];
                    } else if (depth == 16) {
                        s[k] = cur[(static_cast<size_t>(x) * img_n + k) * 2];   // high byte
                    } else {
                        const size_t bo = static_cast<size_t>(x) * depth;
                        const int v = (cur[bo >> 3] >> (8 - depth - static_cast<int>(bo & 7))) & ((1 << depth) - 1);
                        s[k] = static_cast<uint8_t>(v * gray_scale);
                    }
                }
                uint8_t* dst = &out->data[(static_cast<size_t>(y0 + y * dy) * w + (x0 + x * dx)) * color_n];
                if (color == 3) {
                    if (3u * s[0] + 2u >= pal.size()) return bad("palette index out of range");
                    dst[0] = pal[3 * s[0]];
                    dst[1] = pal[3 * s[0] + 1];
                    dst[2] = pal[3 * s[0] + 2];
                } else {
                    for (int k = 0; k < color_n; ++k) dst[k] = s[k];
                }
            }
            std::swap(prev, cur);
        }
    }
    return true;
}

bool decode_tga(const std::vector<uint8_t>& f, Picture8* out, std::string* err) {
    auto bad = [&](const std::string& m) {
        *err = "TGA: " + m;
        return false;
    };
    if (f.size() < 18) return bad("truncated header");
    const int idlen = f[0], cmtype = f[1], itype = f[2];
    const int w = f[12] | (f[13] << 8), h = f[14] | (f[15] << 8), bpp = f[16], desc = f[17];
    if (cmtype != 0 || (itype != 2 && itype != 3 && itype != 10 && itype != 11))
        return bad("only true-colour and gray images (types 2, 3, 10, 11) without a colour map are read");
    const bool gray = itype == 3 || itype == 11, rle = itype >= 9;
    if (gray ? (bpp != 8 && bpp != 16) : (bpp != 24 && bpp != 32)) return bad("unsupported bits per pixel");
    if (w == 0 || h == 0) return bad("empty image");
    const int bytes = bpp / 8;
    {   // the smallest encoding of w x h pixels must fit in the file before anything is allocated
        const size_t npx = static_cast<size_t>(w) * h, avail = f.size() - std::min(f.size(), 18 + size_t(idlen));
        const size_t least = rle ? (npx + 127) / 128 * (1 + static_cast<size_t>(bytes)) : npx * bytes;
        if (least > avail) return bad("image larger than the file holds");
    }
    out->width = w;
    out->height = h;
    out->channels = gray ? 1 : 3;
    out->has_alpha = bytes == 2 || bytes == 4;
    out->data.assign(static_cast<size_t>(w) * h * out->channels, 0);
    size_t pos = 18 + static_cast<size_t>(idlen);
    const size_t npix = static_cast<size_t>(w) * h;
    std::vector<uint8_t> px(npix * bytes);
    if (!rle) {
        if (f.size() - std::min(f.size(), pos) < px.size()) return bad("truncated pixel data");
        std::memcpy(px.data(), &f[pos], px.size());
    } else {
        size_t i = 0;
        while (i < npix) {
            if (pos >= f.size()) return bad("truncated run-length data");
            const int hdr = f[pos++];
            const size_t n = static_cast<size_t>(hdr & 127) + 1;
            if (i + n > npix) return bad("run-length packet overflows the image");
            if (hdr & 128) {
                if (pos + bytes > f.size()) return bad("truncated run-length data");
                for (size_t k = 0; k < n; ++k) std::memcpy(&px[(i + k) * bytes], &f[pos], bytes);
                pos += bytes;
            } else {
                if (pos + n * bytes > f.size()) return bad("truncated run-length data");
                std::memcpy(&px[i * bytes], &f[pos], n * bytes);
                pos += n * bytes;
            }
            i += n;
        }
    }
    // rows are stored bottom-up unless the descriptor's top-left bit is set
    const bool top_down = (desc & 0x20) != 0;
    for (int y = 0; y < h; ++y) {
        const int sy = top_down ? y : h - 1 - y;
        for (int x = 0; x < w; ++x) {
            const uint8_t* s = &px[(static_cast<size_t>(sy) * w + x) * bytes];
            uint8_t* d = &out->data[(static_cast<size_t>(y) * w + x) * out->channels];
            if (gray) {
                d[0] = s[0];
            } else {   // stored B, G, R
                d[0] = s[2];
                d[1] = s[1];
                d[2] = s[0];
            }
        }
    }
    return true;
}

namespace {
bool ends_with(const std::string& path, const char* e) {
    const size_t n = std::strlen(e);
    return path.size() >= n && path.compare(path.size() - n, n, e) == 0;
}
}  // namespace

bool rgb_image_supported(const std::string& path) {
    for (const char* e : {".png", ".PNG", ".tga", ".TGA", ".hdr", ".HDR"})
        if (ends_with(path, e)) return true;
    return false;
}

bool load_rgb_image(const std::string& path, int32_t* width, int32_t* height, std::vector<float>* rgb,
                    std::string* err) {
    auto ends = [&](const char* e) { return ends_with(path, e); };
    if (ends(".hdr") || ends(".HDR")) {
        std::vector<float> raw;
        int w = 0, h = 0;
        if (!read_hdr(path, &raw, &w, &h, err)) return false;
        *width = w;
        *height = h;
        rgb->resize(raw.size());
        for (int y = 0; y < h; ++y)
            std::memcpy(&(*rgb)[static_cast<size_t>(y) * w * 3], &raw[static_cast<size_t>(h - 1 - y) * w * 3],
                        sizeof(float) * 3 * static_cast<size_t>(w));
        return true;
    }
    const bool png = ends(".png") || ends(".PNG"), tga = ends(".tga") || ends(".TGA");
    if (!png && !tga) {
        *err = "image <" + path + ">: only .png, .tga and .hdr pictures are read";
        return false;
    }
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        *err = "cannot open " + path;
        return false;
    }
    const std::vector<uint8_t> file((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    Picture8 pic;
    if (!(png ? decode_png(file, &pic, err) : decode_tga(file, &pic, err))) {
        *err = path + ": " + *err;
        return false;
    }
    *width = pic.width;
    *height = pic.height;
    const size_t npix = static_cast<size_t>(pic.width) * pic.height;
    rgb->assign(npix * 3, 0.0f);
    if (pic.has_alpha) return true;   // the reference keeps such a picture all zero
    // each byte through float sRGB decoding and back to a byte, then / 255
    float lin[256];
    for (int v = 0; v < 256; ++v) {
        const float x = static_cast<float>(v) / 255.0f;
        float l = x <= 0.04045f ? x * 1.0f / 12.92f : std::pow((x + 0.055f) / 1.055f, 2.4f);
        l = l * 255.0f + 0.5f;
        l = l < 0.0f ? 0.0f : (l > 255.0f ? 255.0f : l);
        lin[v] = static_cast<float>(static_cast<uint8_t>(l)) / 255.0f;
    }
    for (int y = 0; y < pic.height; ++y) {
        const int sy = pic.height - 1 - y;
        for (int x = 0; x < pic.width; ++x) {
            const uint8_t* s = &pic.data[(static_cast<size_t>(sy) * pic.width + x) * pic.channels];
            float* d = &(*rgb)[(static_cast<size_t>(y) * pic.width + x) * 3];
            for (int c = 0; c < 3; ++c) d[c] = lin[s[pic.channels == 1 ? 0 : c]];
        }
    }
    return true;
}

}  // namespace mpt
