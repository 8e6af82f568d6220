"""Film filters and the 8-bit film save, on the CPU.

Golden vectors (tests/golden/make_golden.py, written by the compiled reference):
  filters.npz  the sampling renderer's linear film of a Cornell box under every
               filter kind and key set (MitchellFilter, mitchellFilter.cpp:16-40;
               ConeFilter, coneFilter.cpp:13-16; GaussianFilter with explicit
               sigma / amplitude, gaussianFilter.cpp:9-33; a wide box), including
               8 x 8 tiles (footprint clamp, filmTile.cpp:36-50);
  srgb8.npz    Film::save (film.cpp:159-196) of a crafted film (the floats around
               every byte boundary, zero / negative / large weights) and of a
               rendered film, written by the reference as .ppm, .png and .jpg
               (PictureSaver, pictureSaver.cpp:12-82).
Bars: the oracle's films and bytes bit-exact; this library's PPM file
byte-identical to the reference's; its PNG decodes to the same pixels; its
JPEG carries the reference's quantisation and Huffman tables and decodes to
within 1 level of the reference's JPEG (the entropy-coded data is not
byte-compared: parity unpinned at the JPEG encoder's DCT rounding).
"""
import os

import numpy as np
import pytest

from oracle.binding import OracleScene, film_to_srgb8, srgb8_thresholds
from tests.golden.make_golden import filter_scenes
from tests.pictures import decode_jpeg, jpeg_tables, read_png_rgb8, read_ppm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILTERS = dict(filter_scenes())


@pytest.fixture(scope="module")
def films():
    return np.load(os.path.join(GOLDEN, "filters.npz"))


@pytest.fixture(scope="module")
def srgb():
    return np.load(os.path.join(GOLDEN, "srgb8.npz"))


@pytest.mark.parametrize("name", sorted(FILTERS))
def test_oracle_filter_film(films, name):
    f = OracleScene(FILTERS[name]).render_linear(31, threads=4)
    assert np.array_equal(f.view(np.int32), films[name].view(np.int32)), name


def test_filter_keys_change_the_film(films):
    """The keyed fixtures differ from the defaults (so a reader that ignores b, c,
    sigma, amplitude or the filter size fails the test above)."""
    import copy
    for name in ("mitchell_keys", "gaussian_keys", "cone_keys", "box_wide"):
        spec = copy.deepcopy(FILTERS[name])
        spec.filter_params, spec.filter_width, spec.filter_height = {}, None, None
        f = OracleScene(spec).render_linear(31, threads=4)
        assert not np.array_equal(f, films[name]), name


@pytest.mark.parametrize("key", ["crafted", "render"])
def test_oracle_srgb8_bytes(srgb, key):
    hdr, px = read_ppm(srgb[key + "_ppm"].tobytes())
    assert hdr == b"P6 64 48 255\n"
    assert np.array_equal(film_to_srgb8(srgb[key]), px)


def test_srgb8_threshold_table_closed_form():
    """The threshold table against the byte function solved in double precision:
    byte >= k  <=>  s(v) * 255 + 0.5 >= k with s the sRGB OETF (12.92 v below
    0.0031308, 1.055 v^(1/2.4) - 0.055 above). Each float threshold lies within 8
    ulps of the exact solution (measured: at most 4; the reference's float
    arithmetic rounds four times and the subtraction of 0.055 cancels up to ~2.4x);
    a wrong constant or branch moves it by thousands."""
    t = srgb8_thresholds()
    assert t[0] == 0 and np.all(np.diff(t.astype(np.int64)) > 0)
    k = np.arange(1, 256, dtype=np.float64)
    s = (k - 0.5) / 255.0
    v = np.where(s <= 12.92 * 0.0031308, s / 12.92, ((s + 0.055) / 1.055) ** 2.4)
    exact_bits = v.astype(np.float32).view(np.uint32).astype(np.int64)
    d = t[1:].astype(np.int64) - exact_bits
    assert np.abs(d).max() <= 8, (np.abs(d).max(), np.argmax(np.abs(d)) + 1)
    # the scan's own bytes at each threshold and one ulp below
    below = (t[1:] - 1).view(np.float32)
    at = t[1:].view(np.float32)
    film = np.zeros((2, 255, 4), np.float32)
    film[0, :, :3] = below[:, None]
    film[1, :, :3] = at[:, None]
    film[..., 3] = 1.0
    b = film_to_srgb8(film)[..., 0].astype(int)
    assert np.array_equal(b[1], np.arange(1, 256)) and np.array_equal(b[0], np.arange(0, 255))


@pytest.mark.parametrize("key", ["crafted", "render"])
def test_reference_png_holds_the_ppm_pixels(srgb, key):
    _, px = read_ppm(srgb[key + "_ppm"].tobytes())
    assert np.array_equal(read_png_rgb8(srgb[key + "_png"].tobytes()), px)


@pytest.mark.parametrize("key", ["crafted", "render"])
def test_image_write_matches_reference_files(srgb, key, tmp_path):
    from cadise_amd.api import image_write
    ref_ppm = srgb[key + "_ppm"].tobytes()
    _, px = read_ppm(ref_ppm)
    for ext in ("ppm", "png", "jpg"):
        image_write(str(tmp_path / f"o.{ext}"), px)
    assert (tmp_path / "o.ppm").read_bytes() == ref_ppm
    assert np.array_equal(read_png_rgb8((tmp_path / "o.png").read_bytes()), px)
    ours, ref = (tmp_path / "o.jpg").read_bytes(), srgb[key + "_jpg"].tobytes()
    to, tr = jpeg_tables(ours), jpeg_tables(ref)
    assert to["dqt"] == tr["dqt"] and to["dht"] == tr["dht"] and to["sof"] == tr["sof"]
    a, b = decode_jpeg(ours).astype(int), decode_jpeg(ref).astype(int)
    assert np.abs(a - b).max() <= 1 and np.mean(a != b) < 0.01, (np.abs(a - b).max(), np.mean(a != b))
    mse = np.mean((a - px.astype(int)) ** 2)
    assert 10 * np.log10(255.0 ** 2 / mse) > 30.0


def test_image_write_rejects_unknown_extension(tmp_path):
    from cadise_amd.api import image_write
    from cadise_amd.native import MptError
    with pytest.raises(MptError):
        image_write(str(tmp_path / "o.tif"), np.zeros((4, 4, 3), np.uint8))


def test_output_filename_and_host_save(tmp_path):
    """The film block's output-filename (default cadise.jpg,
    filmInstantiator.cpp:24) through the ABI; a host-only context cannot save."""
    from cadise_amd import scenes
    from cadise_amd.api import Context, Scene
    from cadise_amd.native import MptError
    ctx = Context(-1)
    spec = scenes.cornell_box(width=8, height=8, spp=1)
    spec.output = str(tmp_path / "picture.png")
    sc = Scene.from_crsd_text(ctx, spec.to_crsd())
    assert sc.output_filename() == spec.output
    text = spec.to_crsd().replace(f'    [output-filename string "{spec.output}"]\n', "")
    assert "output-filename" not in text
    assert Scene.from_crsd_text(ctx, text).output_filename() == "cadise.jpg"
    with pytest.raises(MptError):
        sc.save()


def test_cornell_reference_image_fixture():
    """tests/golden/cornell_r16k.npz is the reference renderer's 16384-spp image of
    the Cornell box workload (the R of its time-to-relMSE): the oracle's own
    64-spp render of that workload lies within Monte Carlo noise of it (relMSE
    ~3e-3 expected at 64 spp), and a render of another scene does not."""
    from cadise_amd import scenes
    g = np.load(os.path.join(GOLDEN, "cornell_r16k.npz"))
    R = g["image"].astype(np.float64)
    assert R.shape == (256, 256, 3) and int(g["spp"]) == 16384 and np.all(np.isfinite(R))

    def rel(spec):
        f = OracleScene(spec).render_linear(9, threads=8)
        img = f[..., :3] / f[..., 3:]
        return float(np.mean((img - R) ** 2 / (R ** 2 + 0.01)))

    e = rel(scenes.cornell_box(width=256, height=256, spp=64, max_depth=8))
    other = scenes.cornell_box(width=256, height=256, spp=64, max_depth=2)
    e2 = rel(other)
    print(f"relMSE of a 64-spp oracle render vs R: {e:.2e}; depth-2 render: {e2:.2e}")
    assert e < 1e-2 and e2 > 3 * e
