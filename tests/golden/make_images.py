"""Writes the picture fixtures under tests/golden/images/ and, where the compiled
reference is built (oracle/_ref), the texels its PictureLoader::loadRgbImage
reads from each (tests/golden/images.npz).

The pictures are small seeded images written here with the standard library
(zlib, struct), chosen to cover what an image texture may be: PNG colour types
0/2/3/4/6 at every legal bit depth, each row filter, Adam7 interlacing, stored /
fixed / dynamic deflate blocks split over several IDAT chunks, tRNS, ancillary
chunks; TGA types 2/3/10/11 at both row orders; Radiance .hdr flat and
run-length scanlines.

    python tests/golden/make_images.py
"""
from __future__ import annotations

import os
import struct
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "images")


# ---- PNG ----------------------------------------------------------------------------

def _chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _filter_row(ft: int, row: bytes, prev: bytes, bpp: int) -> bytes:
    out = bytearray(len(row))
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        if ft == 0:
            p = 0
        elif ft == 1:
            p = a
        elif ft == 2:
            p = b
        elif ft == 3:
            p = (a + b) >> 1
        else:
            pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
            p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
        out[i] = (x - p) & 255
    return bytes(out)


def _pack_rows(samples: np.ndarray, depth: int) -> list[bytes]:
    """samples: H x W x N integer samples -> packed scanlines (big-endian 16-bit, MSB-first sub-byte)."""
    rows = []
    for r in samples:
        flat = r.reshape(-1)
        if depth == 16:
            rows.append(b"".join(struct.pack(">H", int(v)) for v in flat))
        elif depth == 8:
            rows.append(bytes(int(v) for v in flat))
        else:
            bits = "".join(format(int(v), f"0{depth}b") for v in flat)
            bits += "0" * (-len(bits) % 8)
            rows.append(bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8)))
    return rows


def write_png(path, samples, depth, color, palette=None, trns=None, interlace=False, level=6, strategy=None,
              filters=(0, 1, 2, 3, 4), idat_split=3, extra=()):
    samples = np.asarray(samples)
    h, w = samples.shape[:2]
    n = samples.shape[2] if samples.ndim == 3 else 1
    samples = samples.reshape(h, w, n)
    bpp = max(1, n * depth // 8)
    raw = bytearray()
    passes = ([(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
              if interlace else [(0, 0, 1, 1)])
    k = 0
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        rows = _pack_rows(sub, depth)
        prev = bytes(len(rows[0]))
        for r in rows:
            ft = filters[k % len(filters)]
            k += 1
            raw += bytes([ft]) + _filter_row(ft, r, prev, bpp)
            prev = r
    if strategy is None:
        z = zlib.compress(bytes(raw), level)
    else:
        co = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy)
        z = co.compress(bytes(raw)) + co.flush()
    out = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, int(interlace)))
    for kind, data in extra:
        out += _chunk(kind, data)
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        out += _chunk(b"tRNS", trns)
    step = max(1, -(-len(z) // idat_split))
    for i in range(0, len(z), step):
        out += _chunk(b"IDAT", z[i:i + step])
    out += _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(out)


# ---- TGA -----------------------------------------------------------------------------

def write_tga(path, pix, itype, top_down=False):
    """pix: H x W (gray) or H x W x 3/4 (RGB(A)) uint8, top row first."""
    pix = np.asarray(pix, np.uint8)
    h, w = pix.shape[:2]
    gray = pix.ndim == 2
    bpp = 8 if gray else 8 * pix.shape[2]
    rows = pix if top_down else pix[::-1]
    if gray:
        px = [bytes([int(v)]) for v in rows.reshape(-1)]
    else:
        px = [bytes(p[[2, 1, 0] + ([3] if pix.shape[2] == 4 else [])]) for p in rows.reshape(-1, pix.shape[2])]
    body = bytearray()
    if itype in (2, 3):
        body = b"".join(px)
    else:   # run-length packets: runs of equal pixels, raw packets otherwise (at most 128 each)
        i = 0
        while i < len(px):
            j = i
            while j + 1 < len(px) and px[j + 1] == px[i] and j - i < 127:
                j += 1
            if j > i:
                body += bytes([128 | (j - i)]) + px[i]
                i = j + 1
                continue
            j = i
            while j + 1 < len(px) and px[j + 1] != px[j] and j - i < 127:
                j += 1
            body += bytes([j - i]) + b"".join(px[i:j + 1])
            i = j + 1
    head = struct.pack("<BBBHHBHHHHBB", 3, 0, itype, 0, 0, 0, 0, 0, w, h, bpp, 0x20 if top_down else 0)
    with open(path, "wb") as f:
        f.write(head + b"abc" + bytes(body))


# ---- Radiance .hdr ---------------------------------------------------------------------

def _rgbe(rgb):
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    m = rgb.max(-1)
    nz = m > 1e-32
    e = np.zeros_like(m, dtype=np.int32)
    f = np.zeros_like(m)
    f[nz], e[nz] = np.frexp(m[nz])
    scale = np.where(nz, f * 256.0 / np.where(nz, m, 1.0), 0.0)
    out[..., :3] = np.clip(rgb * scale[..., None], 0, 255).astype(np.uint8)
    out[..., 3] = np.where(nz, e + 128, 0).astype(np.uint8)
    return out


def write_hdr(path, rgb, rle):
    h, w = rgb.shape[:2]
    px = _rgbe(rgb)
    body = bytearray()
    for row in px:
        if not rle:
            body += bytes(row.reshape(-1))
            continue
        body += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            ch = row[:, c]
            i = 0
            while i < w:
                j = i
                while j + 1 < w and ch[j + 1] == ch[i] and j - i < 126:
                    j += 1
                if j - i >= 2:
                    body += bytes([128 + j - i + 1, int(ch[i])])
                    i = j + 1
                else:
                    n = min(128, w - i)
                    body += bytes([n]) + bytes(int(v) for v in ch[i:i + n])
                    i += n
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode() + bytes(body))


def write_all():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(11)
    H, W = 9, 13
    yy, xx = np.mgrid[0:H, 0:W]
    rgb = np.stack([(xx * 19 + yy * 7) % 256, (yy * 31) % 256, rng.integers(0, 256, (H, W))], -1)
    gray = (xx * 20 + yy * 3) % 256
    files = {}

    def png(name, *a, **k):
        write_png(os.path.join(OUT, name), *a, **k)
        files[name] = name

    png("rgb8_filters.png", rgb, 8, 2)
    png("rgb8_stored.png", rgb, 8, 2, level=0, filters=(4,))
    png("rgb8_fixed.png", rgb, 8, 2, level=9, strategy=zlib.Z_FIXED, filters=(1, 2))
    png("rgb8_adam7.png", rgb[:7, :11], 8, 2, interlace=True, filters=(3, 4, 0))
    png("rgb16.png", rgb * 256 + rng.integers(0, 256, rgb.shape), 16, 2)
    png("gray8.png", gray, 8, 0, extra=[(b"gAMA", struct.pack(">I", 45455)), (b"tEXt", b"Comment\x00fixture")])
    png("gray16.png", gray * 256 + 77, 16, 0)
    for d in (1, 2, 4):
        png(f"gray{d}.png", (xx + yy) % (1 << d), d, 0)
    png("gray2_adam7.png", (xx * 3 + yy) % 4, 2, 0, interlace=True)
    pal = rng.integers(0, 256, (16, 3))
    png("pal4.png", (xx + 2 * yy) % 16, 4, 3, palette=pal)
    png("pal8.png", (xx * yy) % 16, 8, 3, palette=pal)
    png("pal2.png", (xx + yy) % 4, 2, 3, palette=pal[:4])
    png("rgba8.png", np.concatenate([rgb, np.full((H, W, 1), 200)], -1), 8, 6)
    png("graya8.png", np.stack([gray, gray // 2], -1), 8, 4)
    png("rgb8_trns.png", rgb, 8, 2, trns=struct.pack(">HHH", 1, 2, 3))
    png("pal8_trns.png", (xx * yy) % 16, 8, 3, palette=pal, trns=bytes([255, 0, 128]))

    def tga(name, *a, **k):
        write_tga(os.path.join(OUT, name), *a, **k)
        files[name] = name

    blocky = np.repeat(np.repeat(rgb[::3, ::3], 3, 0), 3, 1)[:H, :W]
    tga("rgb24.tga", rgb, 2)
    tga("rgb24_top.tga", rgb, 2, top_down=True)
    tga("rgb24_rle.tga", blocky, 10)
    tga("gray8.tga", gray, 3)
    tga("gray8_rle.tga", np.repeat(gray[:, ::4], 4, 1)[:, :W], 11, top_down=True)
    tga("rgba32.tga", np.concatenate([rgb, np.full((H, W, 1), 9)], -1), 2)

    hdr = np.exp(rng.normal(0, 2, (6, 40, 3))).astype(np.float32)
    hdr[:, 10:30] = hdr[:, 10:11]
    write_hdr(os.path.join(OUT, "flat.hdr"), hdr[:, :7], rle=False)
    write_hdr(os.path.join(OUT, "rle.hdr"), hdr, rle=True)
    files["flat.hdr"] = "flat.hdr"
    files["rle.hdr"] = "rle.hdr"
    return sorted(files)


def main():
    names = write_all()
    sys.path.insert(0, ROOT)
    from oracle.binding import ref_available, ref_load_rgb_image
    if not ref_available():
        raise SystemExit("oracle/_ref/libcadise_ref.so is not built (make -C oracle ref)")
    out = {n: ref_load_rgb_image(os.path.join(OUT, n)) for n in names}
    np.savez_compressed(os.path.join(HERE, "images.npz"), **out)
    for n in names:
        print(n, out[n].shape, float(out[n].max()) if out[n].size else None)


if __name__ == "__main__":
    main()
