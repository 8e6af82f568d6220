"""Picture and checkerboard textures on the host (no GPU: a host-only context).

Every picture fixture under tests/golden/images/ (PNG colour types, bit depths,
filters, interlacing and deflate block kinds; TGA types and row orders; flat and
run-length Radiance files; pictures with alpha, which the reference keeps all
zero) becomes, through the product's CRSD "image" texture, exactly the texels
the compiled reference's PictureLoader::loadRgbImage reads from it
(tests/golden/images.npz, written by tests/golden/make_images.py). Checkerboard
textures build the reference's texture list and are not texel parameters.
"""
import os
import zlib

import numpy as np
import pytest

from cadise_amd import scenes
from cadise_amd.api import Context, Scene
from cadise_amd.native import MptError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
IMAGES = os.path.join(GOLDEN, "images")
PICS = np.load(os.path.join(GOLDEN, "images.npz"))

SCENE = """film { [image-width int32 "8"] [image-height int32 "8"] [filter string "box"] }
camera { [type string "perspective-pinhole"] [position vector3r "0 0 5"] [direction vector3r "0 0 -1"]
         [up vector3r "0 1 0"] [fov real "40"] }
renderer { [type string "sampling"] [sampler string "stratified"] [sample-number int32 "4"] [estimator string "pt"] }
accelerator { [type string "bvh"] }
light-cluster { [type string "uniform"] }
%s
material { [type string "matte-lambertian"] [name string "m"] [albedo texture "t"] }
primitive { [type string "triangle"] [name string "p"] [bsdf material "m"] [v1 vector3r "-1 -1 0"]
            [v2 vector3r "1 -1 0"] [v3 vector3r "0 1 0"] }
light { [type string "area"] [primitive primitive "p"] [color rgb "1 1 1"] [watt real "10"] }
"""


def image_block(path, name="t", wrap="repeat"):
    return (f'texture-spectrum {{ [type string "image"] [name string "{name}"] [filename string "{path}"] '
            f'[sample-mode string "nearest"] [wrap-mode string "{wrap}"] }}')


@pytest.fixture(scope="module")
def ctx():
    return Context(-1)


@pytest.mark.parametrize("name", sorted(PICS.files))
def test_picture_texels_match_the_reference(ctx, name):
    sc = Scene.from_crsd_text(ctx, SCENE % image_block(os.path.join(IMAGES, name)))
    got, exp = sc.get_texture(0), PICS[name]
    assert got.shape == exp.shape
    assert np.array_equal(got.view(np.int32), exp.view(np.int32))


def test_picture_fixtures_are_not_trivial():
    """The fixtures exercise what they claim: alpha pictures are zero, the others
    are not, and 8-bit pictures went through the byte round trip (k / 255)."""
    for n in PICS.files:
        a = PICS[n]
        if any(k in n for k in ("rgba", "graya", "trns")):
            assert not a.any(), n
            continue
        assert a.any() and a.std() > 0, n
        if not n.endswith(".hdr"):
            k = np.round(a.astype(np.float64) * 255.0).astype(np.float32)
            assert np.array_equal((k / np.float32(255.0)).view(np.int32), a.view(np.int32)), n


def test_picture_errors(ctx, tmp_path):
    jpg = tmp_path / "x.jpg"
    jpg.write_bytes(b"\xff\xd8\xff")
    for path, code in ((str(jpg), -3), (str(tmp_path / "none.png"), -5)):
        with pytest.raises(MptError) as e:
            Scene.from_crsd_text(ctx, SCENE % image_block(path))
        assert e.value.code == code
    # a damaged deflate stream and a truncated file are errors, not crashes
    good = (open(os.path.join(IMAGES, "rgb8_filters.png"), "rb")).read()
    i = good.index(b"IDAT") + 4
    for bad in (good[:i + 6] + bytes(40) + good[i + 46:], good[: len(good) // 2]):
        p = tmp_path / "bad.png"
        p.write_bytes(bad)
        with pytest.raises(MptError):
            Scene.from_crsd_text(ctx, SCENE % image_block(str(p)))
    # the reference leaves real image textures unimplemented
    real = image_block(os.path.join(IMAGES, "gray8.png")).replace("texture-spectrum", "texture-real")
    text = SCENE.replace('[albedo texture "t"]', '[albedo rgb "0.5 0.5 0.5"]') % real
    with pytest.raises(MptError) as e:
        Scene.from_crsd_text(ctx, text)
    assert e.value.code == -3


def test_png_decoder_rejects_what_it_cannot_read(ctx, tmp_path):
    def chunk(k, d):
        return len(d).to_bytes(4, "big") + k + d + (zlib.crc32(k + d) & 0xFFFFFFFF).to_bytes(4, "big")
    sig = b"\x89PNG\r\n\x1a\n"
    raw = zlib.compress(bytes([0, 1, 2, 3]))
    cases = {
        "badsig.png": b"\x89PNX\r\n\x1a\n" + chunk(b"IHDR", bytes(13)),
        "depth.png": sig + chunk(b"IHDR", (1).to_bytes(4, "big") * 2 + bytes([3, 2, 0, 0, 0])) + chunk(b"IEND", b""),
        "critical.png": sig + chunk(b"IHDR", (1).to_bytes(4, "big") * 2 + bytes([8, 0, 0, 0, 0]))
        + chunk(b"ABCD", b"") + chunk(b"IDAT", raw) + chunk(b"IEND", b""),
        "nopal.png": sig + chunk(b"IHDR", (1).to_bytes(4, "big") * 2 + bytes([8, 3, 0, 0, 0]))
        + chunk(b"IDAT", zlib.compress(bytes([0, 0]))) + chunk(b"IEND", b""),
    }
    for name, data in cases.items():
        p = tmp_path / name
        p.write_bytes(data)
        with pytest.raises(MptError):
            Scene.from_crsd_text(ctx, SCENE % image_block(str(p)))


def test_checkerboard_texture_list(ctx, tmp_path):
    """A checkerboard's odd and even textures precede it (a constant given in place
    of a name becomes a 1 x 1 texture); it has no texels to set; the C-ABI scene
    description builds the same list as the CRSD reader."""
    spec = scenes.checker_cornell(os.path.join(IMAGES, "rgb8_filters.png"), PICS["rgb8_filters.png"])
    a = Scene.from_crsd_text(ctx, spec.to_crsd(sidecar_dir=str(tmp_path)))
    b = Scene.from_spec(ctx, spec)
    for sc in (a, b):
        n = sc.stats()["ntextures"]
        # pic; back: const; rough: const, const, checker; inner: const, checker; nested: const, checker
        infos = [sc.texture_info(i)[:3] for i in range(n)]
        assert n == 10 and infos[0] == (13, 9, 3)
        assert np.array_equal(sc.get_texture(0), PICS["rgb8_filters.png"])
        with pytest.raises(MptError) as e:
            sc.set_texture(2, np.zeros((1, 1, 3), np.float32))
        assert e.value.code == -3
    assert b.tex_index == {"pic": 0}
    assert [a.texture_info(i) for i in range(10)] == [b.texture_info(i) for i in range(10)]
    for i in range(10):
        assert np.array_equal(a.get_texture(i), b.get_texture(i))


def test_checkerboard_errors(ctx):
    base = SCENE.replace('[albedo texture "t"]', '[albedo texture "c0"]')
    unknown = ['texture-spectrum { [type string "constant"] [name string "t"] [value rgb "1 0 0"] }',
               'texture-spectrum { [type string "checkerboard"] [name string "c0"] [odd-number real "2"] '
               '[even-number real "2"] [odd-texture texture "nope"] [even-texture rgb "0 0 1"] }']
    with pytest.raises(MptError) as e:
        Scene.from_crsd_text(ctx, base % "\n".join(unknown))
    assert e.value.code == -2

    def chain(depth):   # c0 = a checkerboard of c1 ... of a constant, `depth` checkerboards deep
        blocks = [f'texture-spectrum {{ [type string "constant"] [name string "c{depth}"] [value rgb "1 0 0"] }}']
        for k in range(depth - 1, -1, -1):
            blocks.append(f'texture-spectrum {{ [type string "checkerboard"] [name string "c{k}"] '
                          f'[odd-number real "2"] [even-number real "2"] [odd-texture texture "c{k + 1}"] '
                          f'[even-texture rgb "0 0 1"] }}')
        return base % "\n".join(blocks)
    with pytest.raises(MptError) as e:
        Scene.from_crsd_text(ctx, chain(9))
    assert e.value.code == -3
    assert Scene.from_crsd_text(ctx, chain(8)).stats()["ntextures"] == 17
