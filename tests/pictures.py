"""Readers for the 8-bit pictures the film save writes, used to check files
(PPM header and payload, PNG through zlib and the five row filters of the PNG
specification, section 9)."""
import struct
import zlib

import numpy as np


def read_ppm(data: bytes):
    """(header bytes, H x W x 3 uint8) of a binary PPM written as 'P6 w h 255\\n'
    or with any whitespace between the fields."""
    fields, pos = [], 0
    while len(fields) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        start = pos
        while not data[pos:pos + 1].isspace():
            pos += 1
        fields.append(data[start:pos])
    pos += 1   # the single whitespace byte after maxval
    assert fields[0] == b"P6" and fields[3] == b"255", fields
    w, h = int(fields[1]), int(fields[2])
    return data[:pos], np.frombuffer(data[pos:pos + 3 * w * h], np.uint8).reshape(h, w, 3)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def read_png_rgb8(data: bytes) -> np.ndarray:
    """H x W x 3 uint8 of a non-interlaced 8-bit RGB PNG (colour type 2)."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(kind + body) & 0xFFFFFFFF == crc, kind
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            assert (depth, ctype, interlace) == (8, 2, 0), (depth, ctype, interlace)
        elif kind == b"IDAT":
            idat += body
        elif kind == b"IEND":
            break
        pos += 12 + n
    raw = zlib.decompress(idat)
    stride = 3 * w
    out = np.zeros((h, stride), np.uint8)
    prev = [0] * stride
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = list(raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)])
        for i in range(stride):
            a = line[i - 3] if i >= 3 else 0
            b = prev[i]
            c = prev[i - 3] if i >= 3 else 0
            if f == 1:
                line[i] = (line[i] + a) & 255
            elif f == 2:
                line[i] = (line[i] + b) & 255
            elif f == 3:
                line[i] = (line[i] + (a + b) // 2) & 255
            elif f == 4:
                line[i] = (line[i] + _paeth(a, b, c)) & 255
            else:
                assert f == 0, f
        out[y] = line
        prev = line
    return out.reshape(h, w, 3)


# ---- baseline JPEG (ITU-T T.81), for decoding the writer's and the reference's files ----

def _segments(data: bytes):
    """(marker, payload) of every marker segment up to SOS; then ('scan', entropy data)."""
    assert data[:2] == b"\xff\xd8"
    pos = 2
    while pos < len(data):
        assert data[pos] == 0xFF, pos
        m = data[pos + 1]
        pos += 2
        if m == 0xD9:
            return
        n = (data[pos] << 8) | data[pos + 1]
        body = data[pos + 2:pos + n]
        pos += n
        yield m, body
        if m == 0xDA:
            end = pos
            while not (data[end] == 0xFF and data[end + 1] not in (0x00,) and not 0xD0 <= data[end + 1] <= 0xD7):
                end += 1
            yield "scan", data[pos:end]
            pos = end


def jpeg_tables(data: bytes):
    """{'dqt': {id: 64 values in zig-zag order}, 'dht': {(class, id): (bits, vals)},
    'sof': (precision, h, w, [(id, hs, vs, tq)])} of a baseline JPEG."""
    out = {"dqt": {}, "dht": {}, "sof": None}
    for m, body in _segments(data):
        if m == 0xDB:
            p = 0
            while p < len(body):
                pq, tq = body[p] >> 4, body[p] & 15
                assert pq == 0
                out["dqt"][tq] = list(body[p + 1:p + 65])
                p += 65
        elif m == 0xC4:
            p = 0
            while p < len(body):
                tc, th = body[p] >> 4, body[p] & 15
                bits = list(body[p + 1:p + 17])
                vals = list(body[p + 17:p + 17 + sum(bits)])
                out["dht"][(tc, th)] = (bits, vals)
                p += 17 + sum(bits)
        elif m == 0xC0:
            prec, h, w, nc = body[0], (body[1] << 8) | body[2], (body[3] << 8) | body[4], body[5]
            comps = [(body[6 + 3 * i], body[7 + 3 * i] >> 4, body[7 + 3 * i] & 15, body[8 + 3 * i]) for i in range(nc)]
            out["sof"] = (prec, h, w, comps)
        elif m in (0xC1, 0xC2, 0xC3):
            raise ValueError("not a baseline JPEG")
    return out


def _zigzag():
    zz = []
    for s in range(15):
        if s % 2 == 0:
            zz += [8 * r + (s - r) for r in range(min(s, 7), -1, -1) if s - r <= 7]
        else:
            zz += [8 * (s - c) + c for c in range(min(s, 7), -1, -1) if s - c <= 7]
    return zz


def decode_jpeg(data: bytes) -> np.ndarray:
    """H x W x 3 uint8 of a baseline, non-restart, 3-component JPEG (JFIF YCbCr),
    by the T.81 decoding process: Huffman decoding (Annex F.2), dequantisation,
    the inverse DCT of A.3.3 from its definition, nearest upsampling of
    subsampled components, and the JFIF YCbCr -> RGB conversion."""
    t = jpeg_tables(data)
    scan = None
    sos = None
    for m, body in _segments(data):
        if m == 0xDA:
            sos = body
        elif m == "scan":
            scan = body
    _, H, W, comps = t["sof"]
    ns = sos[0]
    sel = {sos[1 + 2 * i]: (sos[2 + 2 * i] >> 4, sos[2 + 2 * i] & 15) for i in range(ns)}
    # Huffman lookup: (length, code) -> symbol
    luts = {}
    for key, (bits, vals) in t["dht"].items():
        lut, code, k = {}, 0, 0
        for ln in range(1, 17):
            for _ in range(bits[ln - 1]):
                lut[(ln, code)] = vals[k]
                code += 1
                k += 1
            code <<= 1
        luts[key] = lut
    raw = scan.replace(b"\xff\x00", b"\xff")
    bitpos = [0]

    def bit():
        i = bitpos[0]
        bitpos[0] += 1
        return (raw[i >> 3] >> (7 - (i & 7))) & 1 if (i >> 3) < len(raw) else 1

    def receive(s):
        v = 0
        for _ in range(s):
            v = (v << 1) | bit()
        return v

    def extend(v, s):
        return v - (1 << s) + 1 if s and v < (1 << (s - 1)) else v

    def huff(lut):
        code, ln = 0, 0
        while True:
            code = (code << 1) | bit()
            ln += 1
            if (ln, code) in lut:
                return lut[(ln, code)]
            assert ln < 16

    zz = _zigzag()
    hmax = max(c[1] for c in comps)
    vmax = max(c[2] for c in comps)
    mcux, mcuy = (W + 8 * hmax - 1) // (8 * hmax), (H + 8 * vmax - 1) // (8 * vmax)
    planes = {c[0]: np.zeros((mcuy * c[2] * 8, mcux * c[1] * 8)) for c in comps}
    pred = {c[0]: 0 for c in comps}
    M = np.array([[(np.sqrt(0.5) if u == 0 else 1.0) * 0.5 * np.cos((2 * x + 1) * u * np.pi / 16) for x in range(8)]
                  for u in range(8)])
    for my in range(mcuy):
        for mx in range(mcux):
            for cid, hs, vs, tq in comps:
                td, ta = sel[cid]
                q = np.zeros(64)
                q[zz] = t["dqt"][tq]
                for by in range(vs):
                    for bx in range(hs):
                        coef = np.zeros(64)
                        s = huff(luts[(0, td)])
                        pred[cid] += extend(receive(s), s)
                        coef[0] = pred[cid]
                        k = 1
                        while k < 64:
                            rs = huff(luts[(1, ta)])
                            r, s = rs >> 4, rs & 15
                            if s == 0:
                                if r == 15:
                                    k += 16
                                    continue
                                break
                            k += r
                            coef[zz[k]] = extend(receive(s), s)
                            k += 1
                        F = (coef * q).reshape(8, 8)
                        f = M.T @ F @ M
                        y0, x0 = (my * vs + by) * 8, (mx * hs + bx) * 8
                        planes[cid][y0:y0 + 8, x0:x0 + 8] = f + 128.0
    ch = []
    for cid, hs, vs, _ in comps:
        p = planes[cid]
        p = np.repeat(np.repeat(p, vmax // vs, axis=0), hmax // hs, axis=1)
        ch.append(p[:H, :W])
    Y, Cb, Cr = ch[0], ch[1] - 128.0, ch[2] - 128.0
    rgb = np.stack([Y + 1.402 * Cr, Y - 0.344136 * Cb - 0.714136 * Cr, Y + 1.772 * Cb], -1)
    return np.clip(np.round(rgb), 0, 255).astype(np.uint8)
