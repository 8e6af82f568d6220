"""GPU parity of the film filters and the 8-bit film save.

  * Every filter kind and key set of the film block (Mitchell, cone, Gaussian
    with explicit sigma / amplitude, a wide box; 8 x 8 tiles), rendered on the
    GPU and by the oracle with the GPU's random layout: per-pixel linear
    radiance within 1e-4 relative (+1e-6) on >= 99% of pixels, image mean
    within 1e-3, filter weight sums within 1e-5 relative (+1e-6).
  * The 8-bit resolve (mpt_film_srgb8_device), bit-exact: on every float in
    [+0, 1] against the oracle's exhaustive threshold table (over 10^9 values:
    the byte is a function of one float, so this covers every weight-1 input),
    on negative, large, infinite and NaN values against the oracle, and on the
    reference's own crafted film against the reference's PPM bytes.
  * mpt_film_save: the PPM file is the reference's header plus the oracle's
    bytes of the GPU film; PNG decodes to them; the JPEG carries the reference's
    tables; NULL writes to the film block's output-filename.
"""
import os

import numpy as np
import pytest

from cadise_amd import scenes
from oracle.binding import OR_RNG_PHILOX, OracleScene, film_to_srgb8, srgb8_thresholds
from tests.golden.make_golden import filter_scenes
from tests.hipmem import DeviceBuffer, synchronize
from tests.pictures import decode_jpeg, jpeg_tables, read_png_rgb8, read_ppm

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    from cadise_amd.api import Context
    return Context(0)


@pytest.mark.parametrize("name", [n for n, _ in filter_scenes()])
def test_filter_film_vs_oracle(ctx, name):
    from cadise_amd.api import Scene, film_mean
    spec = dict(filter_scenes())[name]
    sc = Scene.from_spec(ctx, spec)
    orf = OracleScene(spec).render_linear(99, rng_mode=OR_RNG_PHILOX)
    for passwise in (True, False):
        film = sc.render(seed=99, passwise=passwise)
        g, o = film_mean(film), film_mean(orf)
        close = np.all(np.abs(g - o) <= 1e-4 * np.abs(o) + 1e-6, axis=-1)
        rel = abs(g.mean() - o.mean()) / max(abs(o.mean()), 1e-12)
        print(f"[{name} passwise={passwise}] close={close.mean():.5f} rel={rel:.2e}", flush=True)
        assert close.mean() >= 0.99 and rel < 1e-3, name
        np.testing.assert_allclose(film[..., 3], orf[..., 3], rtol=1e-5, atol=1e-6)


def _srgb8_gpu(ctx, film):
    """mpt_film_srgb8_device of a host film (n x 4 float32) -> n x 3 bytes."""
    from cadise_amd.api import srgb8_device
    film = np.ascontiguousarray(film, np.float32).reshape(-1, 4)
    n = len(film)
    out = np.zeros((n, 3), np.uint8)
    with DeviceBuffer(film.nbytes) as f, DeviceBuffer(out.nbytes) as o:
        f.upload(film)
        srgb8_device(ctx, f.ptr.value, n, o.ptr.value)
        synchronize()
        o.download(out)
    return out


def test_srgb8_every_float_in_unit_interval(ctx):
    t = srgb8_thresholds().astype(np.int64)
    top = 0x3F800000   # 1.0f
    chunk = 3 * (1 << 25)
    mism = 0
    for start in range(0, top + 1, chunk):
        stop = min(top + 1, start + chunk)
        npix = (stop - start + 2) // 3
        bits = np.minimum(np.arange(start, start + 3 * npix, dtype=np.int64), top)
        film = np.ones((npix, 4), np.float32)
        film[:, :3] = bits.astype(np.uint32).view(np.float32).reshape(npix, 3)
        got = _srgb8_gpu(ctx, film).reshape(-1)
        # expected byte of every bit pattern: runs between consecutive thresholds
        edges = np.clip(np.concatenate([t, [top + 1]]), start, start + 3 * npix)
        want = np.repeat(np.arange(256, dtype=np.uint8), np.diff(edges))
        want = np.concatenate([want, np.full(3 * npix - len(want), 255, np.uint8)])   # the clamped tail is 1.0f
        bad = np.flatnonzero(got != want)
        if len(bad):
            i = bad[0]
            print(f"first mismatch at bits {int(bits[i]):#x}: gpu {int(got[i])} oracle {int(want[i])}", flush=True)
        mism += len(bad)
    print(f"[srgb8] floats in [0, 1]: {top + 1}, mismatches {mism}", flush=True)
    assert mism == 0


def test_srgb8_outside_unit_interval_and_weights(ctx):
    rng = np.random.default_rng(3)
    n = 1 << 20
    film = np.zeros((n, 4), np.float32)
    neg = rng.integers(0x80000000, 0xFF800001, (n, 3), dtype=np.uint64).astype(np.uint32).view(np.float32)
    big = rng.integers(0x3F800001, 0x7F800001, (n, 3), dtype=np.uint64).astype(np.uint32).view(np.float32)
    film[:, :3] = np.where(rng.random((n, 1)) < 0.5, neg, big)
    film[:, 3] = 1.0
    film[:8, :3] = np.array([[np.inf, -np.inf, np.nan]] * 8, np.float32)
    # sums with weights: zero, negative, large, subnormal results
    m = n // 2
    w = rng.uniform(-4, 16, m).astype(np.float32)
    w[::7] = 0.0
    film[m:, 3] = w
    film[m:, :3] = (rng.uniform(-0.5, 2.0, (m, 3)) * w[:, None]).astype(np.float32)
    got = _srgb8_gpu(ctx, film)
    want = film_to_srgb8(film)
    assert np.array_equal(got, want), int((got != want).sum())


def test_srgb8_reference_crafted_film(ctx):
    g = np.load(os.path.join(GOLDEN, "srgb8.npz"))
    for key in ("crafted", "render"):
        _, px = read_ppm(g[key + "_ppm"].tobytes())
        got = _srgb8_gpu(ctx, g[key]).reshape(px.shape)
        assert np.array_equal(got, px), (key, int((got != px).sum()))


def test_film_save_files(ctx, tmp_path):
    from cadise_amd.api import Scene
    spec = scenes.cornell_box(width=64, height=48, spp=8, max_depth=6, filter="gaussian")
    spec.output = str(tmp_path / "default_name.ppm")
    sc = Scene.from_spec(ctx, spec)
    film = sc.render(seed=3)
    want = film_to_srgb8(film)
    assert np.array_equal(sc.srgb8(), want)
    for ext in ("ppm", "png", "jpg", "pfm"):
        sc.save(str(tmp_path / f"o.{ext}"))
    sc.save()   # the film block's output-filename
    ppm = (tmp_path / "o.ppm").read_bytes()
    assert ppm == b"P6 64 48 255\n" + want.tobytes()
    assert (tmp_path / "default_name.ppm").read_bytes() == ppm
    assert np.array_equal(read_png_rgb8((tmp_path / "o.png").read_bytes()), want)
    jpg = (tmp_path / "o.jpg").read_bytes()
    ref = np.load(os.path.join(GOLDEN, "srgb8.npz"))["render_jpg"].tobytes()
    assert jpeg_tables(jpg)["dqt"] == jpeg_tables(ref)["dqt"]
    dec = decode_jpeg(jpg).astype(int)
    assert np.sqrt(np.mean((dec - want.astype(int)) ** 2)) < 4.0
    pfm = (tmp_path / "o.pfm").read_bytes()
    head = b"PF\n64 48\n-1.0\n"
    assert pfm.startswith(head)
    lin = np.frombuffer(pfm[len(head):], np.float32).reshape(48, 64, 3)[::-1]
    w = film[..., 3:]
    np.testing.assert_array_equal(lin, film[..., :3] * np.where(w == 0, 0, 1 / np.where(w == 0, 1, w)))
