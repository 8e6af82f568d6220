"""mpt_trace_device_noalloc: the traversal with its work counter, statistics
and stack spill in the library's fixed per-device scratch array instead of
allocated buffers. Through the C ABI its hits must equal mpt_trace_device's
bit for bit (closest hit: primitive and t; any hit: flags), on two calls in a
row with the second one smaller (the array keeps the first call's contents,
which the entry point must clear), on a shallow tree (stacks within LDS) and
on a deeper one (stacks spill into the array), and its argument checks must
reject what the existing entry point rejects."""
import numpy as np
import pytest

from cadise_amd import native, scenes
from tests.helpers import random_rays, small_spheres
from tests.hipmem import DeviceBuffer, synchronize

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cadise_amd.api import Context
    return Context(0)


def _trace(fn, sc, rays, any_hit):
    n = len(rays)
    hits = np.zeros((n, 2), np.int32)
    with DeviceBuffer(rays.nbytes) as r, DeviceBuffer(hits.nbytes) as h:
        r.upload(rays)
        h.upload(np.full((n, 2), -7, np.int32))
        rc = fn(sc.h, r.ptr, n, int(any_hit), h.ptr, None)
        assert rc == 0, (rc, sc.ctx.L.mpt_last_error(sc.ctx.h))
        synchronize()
        h.download(hits)
    return hits


@pytest.mark.parametrize("case", ["spheres", "cornell"])
def test_noalloc_matches_trace_device(ctx, case):
    from cadise_amd.api import Scene
    L = native.load_library()
    if case == "spheres":
        spec = small_spheres(n_spheres=8, level=4)
        lo, hi = (-18, 0, -18), (18, 8, 18)
    else:
        spec = scenes.cornell_box(width=8, height=8, spp=1)
        lo, hi = (-0.1, -0.1, -0.1), (1.1, 1.1, 1.1)
    sc = Scene.from_spec(ctx, spec)
    st = sc.stats()
    stack = 3 * (st["wide_depth"] + 1) + 1   # BVH4: (W - 1)(depth + 1) + 1 entries at most
    print(f"[noalloc {case}] {spec.num_prims()} prims, wide depth {st['wide_depth']}, stack bound {stack} "
          f"({'spills past' if stack > 24 else 'within'} the 24 LDS entries)", flush=True)
    assert (stack > 24) == (case == "spheres")
    big = random_rays(300000, lo, hi, seed=21)
    small = random_rays(1000, lo, hi, seed=22)
    for any_hit in (False, True):
        if any_hit:
            big[:, 7] = np.random.default_rng(23).uniform(0.05, 3.0, len(big)).astype(np.float32)
            small[:, 7] = np.random.default_rng(24).uniform(0.05, 3.0, len(small)).astype(np.float32)
        for rays in (big, small):   # two calls in a row, the second smaller
            want = _trace(L.mpt_trace_device, sc, rays, any_hit)
            got = _trace(L.mpt_trace_device_noalloc, sc, rays, any_hit)
            col = [0] if any_hit else [0, 1]
            assert np.array_equal(got[:, col], want[:, col]), (case, any_hit, len(rays),
                                                              int((got[:, col] != want[:, col]).any(1).sum()))
            hits = int((want[:, 0] == 1).sum()) if any_hit else int((want[:, 0] >= 0).sum())
            print(f"[noalloc {case} any={any_hit}] {len(rays)} rays: {hits} hits, identical", flush=True)


def test_noalloc_argument_checks(ctx):
    from cadise_amd.api import Scene
    L = native.load_library()
    sc = Scene.from_spec(ctx, scenes.cornell_box(width=8, height=8, spp=1))
    cap = (1 << 31) - (1 << 24)
    with DeviceBuffer(64) as b:
        assert L.mpt_trace_device_noalloc(sc.h, b.ptr, cap + 1, 0, b.ptr, None) == native.MPT_ERR_INVALID
        assert L.mpt_trace_device_noalloc(sc.h, None, 5, 0, None, None) == native.MPT_ERR_INVALID
        assert L.mpt_trace_device_noalloc(sc.h, b.ptr, 0, 0, b.ptr, None) == 0
    assert L.mpt_trace_device_noalloc(None, None, 0, 0, None, None) == native.MPT_ERR_INVALID
