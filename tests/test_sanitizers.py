"""CPU builds under AddressSanitizer + UndefinedBehaviorSanitizer
(tests/sanitize/Makefile):

  * the host half of libmpt (CRSD reader, BVH build and wide collapse, the
    PNG / TGA / Radiance decoders, the PPM / PNG / JPEG writers) driven by
    tests/sanitize/host_check over well-formed inputs and seeded malformed
    ones: truncations, byte flips, header fields set to huge or zero values
    (with PNG chunk CRCs recomputed, so the decoder gets past its checksum),
    CRSD text with cut, deleted or replaced characters and truncated binary
    side-car meshes. Every input must be accepted or rejected with an error;
    any sanitizer report fails the test;
  * the oracle (oracle/*.c) loaded under the sanitizer runtime while the
    oracle's golden-vector tests run.
"""
import os
import struct
import subprocess
import sys
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
BIN = os.path.join(SAN, "_build", "host_check")
IMAGES = os.path.join(ROOT, "tests", "golden", "images")
MARKERS = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer", "SUMMARY: UndefinedBehaviorSanitizer")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", SAN], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + r.stderr[-400:])
    return BIN


def _run(args, timeout=300, env=None):
    r = subprocess.run(args, capture_output=True, text=True, errors="replace", timeout=timeout, env=env)
    out = r.stdout + r.stderr
    bad = [m for m in MARKERS if m in out]
    assert r.returncode == 0 and not bad, (args[:3], r.returncode, bad, out[-3000:])
    return out


def _png_chunks(data):
    pos, out = 8, []
    while pos + 8 <= len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        out.append((data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]))
        pos += 12 + n
    return out


def _png_from(chunks):
    out = b"\x89PNG\r\n\x1a\n"
    for kind, body in chunks:
        out += struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)
    return out


def _mutations(data, rng, ext):
    yield data
    for cut in sorted({0, 1, 8, 12, 18, 33, len(data) // 2, len(data) - 1}):
        if cut < len(data):
            yield data[:cut]
    for _ in range(24):
        b = bytearray(data)
        for _ in range(int(rng.integers(1, 6))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        yield bytes(b)
    if ext == ".png":
        ch = _png_chunks(data)
        for w, h in ((0, 4), (1 << 30, 1 << 30), (0xFFFFFFFF, 1), (65536, 65536), (3, 0x7FFFFFFF)):
            c2 = [(k, struct.pack(">II", w, h) + b[8:] if k == b"IHDR" else b) for k, b in ch]
            yield _png_from(c2)
        for depth, ctype in ((16, 3), (3, 2), (8, 5), (1, 6)):
            c2 = [(k, b[:8] + bytes([depth, ctype]) + b[10:] if k == b"IHDR" else b) for k, b in ch]
            yield _png_from(c2)
        # the image data deflated again, cut short or with its stream damaged
        idat = b"".join(b for k, b in ch if k == b"IDAT")
        raw = zlib.decompress(idat)
        for body in (zlib.compress(raw[: len(raw) // 2]), zlib.compress(raw + raw), idat[:-4], idat[:7]):
            yield _png_from([(k, b) for k, b in ch if k not in (b"IDAT", b"IEND")] + [(b"IDAT", body), (b"IEND", b"")])
    if ext == ".tga" and len(data) >= 18:
        for w, h in ((0, 5), (65535, 65535), (1, 65535)):
            b = bytearray(data)
            b[12:16] = struct.pack("<HH", w, h)
            yield bytes(b)
    if ext == ".hdr":
        text = data.split(b"\n")
        for dims in (b"-Y 1000000 +X 1000000", b"-Y 0 +X 4", b"-Y -3 +X 2", b"+Y 2 -X 2", b"-Y 2"):
            yield b"\n".join(dims if t.startswith((b"-Y", b"+Y")) else t for t in text)


def test_picture_decoders_under_sanitizers(built, tmp_path):
    rng = np.random.default_rng(17)
    files = {"png": [], "tga": [], "image": []}
    n = 0
    for name in sorted(os.listdir(IMAGES)):
        ext = os.path.splitext(name)[1]
        data = open(os.path.join(IMAGES, name), "rb").read()
        for k, m in enumerate(_mutations(data, rng, ext)):
            p = tmp_path / f"{n:05d}{ext}"
            p.write_bytes(m)
            n += 1
            files["image"].append(str(p))
            if ext in (".png", ".tga"):
                files[ext[1:]].append(str(p))
    out = ""
    for kind, paths in files.items():
        for i in range(0, len(paths), 400):
            out += _run([built, kind] + paths[i:i + 400])
    acc, rej = out.count("accepted "), out.count("rejected ")
    print(f"{n} picture files: {acc} decodes accepted, {rej} rejected")
    assert acc > 50 and rej > 50


def _crsd_mutations(text, rng):
    yield text
    for frac in (0.0, 0.1, 0.5, 0.9, 0.999):
        yield text[: int(len(text) * frac)]
    for _ in range(16):
        t = list(text)
        for _ in range(int(rng.integers(1, 8))):
            i = int(rng.integers(0, len(t)))
            op = int(rng.integers(0, 3))
            if op == 0:
                del t[i]
            elif op == 1:
                alphabet = "{}[]\"\n#0-e9 "
                t[i] = alphabet[int(rng.integers(0, len(alphabet)))]
            else:
                t.insert(i, "]")
        yield "".join(t)
    for a, b in (('"256"', '"-7"'), ('"256"', '"99999999999"'), ('"16"', '"0"'), ('"8"', '"-1"'),
                 ("0.63", "1e40"), ("0.63", "nan"), ('"box"', '"lanczos"'), ('"sah"', '"middle"'),
                 ("vector3r", "vector3r-array"), ("triangle", "rectangle"), ("rgb", "real")):
        yield text.replace(a, b)


def test_crsd_reader_and_bvh_under_sanitizers(built, tmp_path):
    from cadise_amd import scenes
    rng = np.random.default_rng(23)
    specs = [scenes.cornell_box(width=32, height=32, spp=2), scenes.shipped_cornell(width=32, height=24, spp=2),
             scenes.material_zoo(width=24, height=16, spp=2),
             scenes.procedural_spheres(n_spheres=3, level=2, width=24, height=16, spp=1, offset=(0, 0, 0)),
             scenes.sky_spheres(n_spheres=2, level=1, width=16, height=12, spp=1, env_size=(16, 8),
                                sun_diameter_deg=10.0, offset=(0, 0, 0))]
    paths = []
    for si, spec in enumerate(specs):
        d = tmp_path / f"s{si}"
        d.mkdir()
        text = spec.to_crsd(sidecar_dir=str(d), sidecar_threshold=8)
        for k, t in enumerate(_crsd_mutations(text, rng)):
            p = d / f"m{k:03d}.crsd"
            p.write_text(t)
            paths.append(str(p))
        # truncated and empty binary side-cars behind the unmodified text
        for f in sorted(os.listdir(d)):
            if f.endswith(".f32"):
                data = (d / f).read_bytes()
                (d / f).write_bytes(data[: len(data) // 3 * 2 + 1])
                break
        p = d / "sidecar_cut.crsd"
        p.write_text(text)
        paths.append(str(p))
    out = _run([built, "crsd"] + paths, timeout=600)
    acc, rej = out.count("accepted "), out.count("rejected ")
    print(f"{len(paths)} CRSD files: {acc} accepted, {rej} rejected")
    assert acc >= len(specs) and rej > 20


def test_picture_writers_under_sanitizers(built, tmp_path):
    assert "writers ok" in _run([built, "write", str(tmp_path)])
    from tests.pictures import decode_jpeg, read_png_rgb8, read_ppm
    for w, h in ((7, 5), (64, 17)):
        _, a = read_ppm((tmp_path / f"w{w}x{h}.ppm").read_bytes())
        assert np.array_equal(read_png_rgb8((tmp_path / f"w{w}x{h}.png").read_bytes()), a)
        assert decode_jpeg((tmp_path / f"w{w}x{h}.jpg").read_bytes()).shape == a.shape


def test_oracle_under_sanitizers(built):
    lib = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.exists(lib):
        pytest.skip("libasan.so not found")
    env = dict(os.environ, LD_PRELOAD=lib, MPT_ORACLE_SO=os.path.join(SAN, "_build", "liboracle.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    out = _run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                os.path.join(ROOT, "tests", "test_oracle_golden.py"),
                os.path.join(ROOT, "tests", "test_film_save.py") + "::test_oracle_filter_film",
                os.path.join(ROOT, "tests", "test_film_save.py") + "::test_srgb8_threshold_table_closed_form"],
               timeout=900, env=env)
    assert " passed" in out and " failed" not in out
