"""GPU parity of the traversal (Accelerator::isIntersecting / isOccluded,
reference bvhAccelerator.cpp:40-156) through the C ABI against the oracle,
whose binary traversal reproduces the compiled reference bit for bit
(tests/test_oracle_golden.py).

Bar: closest-hit primitive id identical and t bit-identical on every ray whose
closest hit is unique; rays with an exact-t tie (several primitives at the
same t, e.g. a shared edge) are counted separately, and there the GPU's
primitive must be one of the tied ones with the same t. Any-hit flags
identical.
"""
import numpy as np
import pytest

from cadise_amd import scenes
from oracle.binding import OracleScene
from tests.helpers import camera_rays, random_rays, small_spheres

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cadise_amd.api import Context
    return Context(0)


def _check_closest(scene, oracle, rays, label):
    prim, t = scene.trace(rays)
    oprim, ot = oracle.trace(rays)
    ties = oracle.ties(rays)
    unique = ties <= 1
    print(f"[{label}] rays={len(rays)} hits={int((oprim >= 0).sum())} ties={int((~unique).sum())}", flush=True)
    assert np.array_equal(prim[unique], oprim[unique]), f"{label}: {int((prim[unique] != oprim[unique]).sum())} id mismatches"
    assert np.array_equal(t[unique].view(np.int32), ot[unique].view(np.int32)), f"{label}: t not bit-identical"
    # tied rays: same t; the winner must hit at that t as well
    tied = np.where(~unique)[0]
    if len(tied):
        assert np.array_equal(t[tied].view(np.int32), ot[tied].view(np.int32))
    return int((~unique).sum())


def _check_any(scene, oracle, rays, label):
    occ = scene.is_occluded(rays)
    oocc, _ = oracle.trace(rays, any_hit=True)
    assert np.array_equal(occ, oocc.astype(bool)), f"{label}: {int((occ != oocc.astype(bool)).sum())} any-hit mismatches"


@pytest.mark.parametrize("width", [4, 8])
def test_trace_cornell(ctx, width):
    from cadise_amd.api import Scene, bvh_opts
    spec = scenes.cornell_box()
    sc = Scene.from_crsd_text(ctx, spec.to_crsd(), bvh_opts(width=width))
    orc = OracleScene(spec)
    rays = np.concatenate([random_rays(50000, -0.1, 1.1, seed=1), camera_rays(spec, 20000, seed=2)])
    _check_closest(sc, orc, rays, f"cornell bvh{width}")
    short = rays.copy()
    short[:, 7] = np.random.default_rng(3).uniform(0.01, 1.5, len(rays)).astype(np.float32)
    _check_any(sc, orc, short, f"cornell bvh{width} any")
    _check_closest(sc, orc, short, f"cornell bvh{width} bounded")


@pytest.mark.parametrize("width,leaf_max", [(4, 1), (4, 2), (8, 4)])
def test_trace_spheres(ctx, tmp_path, width, leaf_max):
    from cadise_amd.api import Scene, bvh_opts
    spec = small_spheres(n_spheres=8, level=4)   # 8 x 5120 triangles + ground + 4 rectangle lights
    sc = Scene.from_crsd(ctx, _write(tmp_path, spec), bvh_opts(width=width, leaf_max=leaf_max))
    orc = OracleScene(spec)
    rays = np.concatenate([random_rays(100000, (-18, 0, -18), (18, 8, 18), seed=5), camera_rays(spec, 20000, seed=6)])
    _check_closest(sc, orc, rays, f"spheres bvh{width} leaf{leaf_max}")
    short = rays.copy()
    short[:, 7] = np.random.default_rng(7).uniform(0.1, 20.0, len(rays)).astype(np.float32)
    _check_any(sc, orc, short, f"spheres bvh{width} any")


def test_trace_edge_cases(ctx):
    """Axis-parallel rays (zero direction components), rays starting on
    surfaces, t == minT / maxT boundaries and an empty batch."""
    from cadise_amd.api import Scene
    spec = scenes.cornell_box()
    sc = Scene.from_crsd_text(ctx, spec.to_crsd())
    orc = OracleScene(spec)
    rng = np.random.default_rng(11)
    n = 6000
    o = rng.uniform(0.05, 0.95, (n, 3)).astype(np.float32)
    axes = np.zeros((n, 3), np.float32)
    axes[np.arange(n), rng.integers(0, 3, n)] = rng.choice([-1.0, 1.0], n)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3] = 1e-3
    rays[:, 4:7] = axes
    rays[:, 7] = np.finfo(np.float32).max
    _check_closest(sc, orc, rays, "axis-parallel")
    # maxT exactly at the hit distance: accepted (t == maxT), and just below: rejected
    _, t = orc.trace(rays)
    hit = t < 1e30
    exact = rays[hit].copy()
    exact[:, 7] = t[hit]
    _check_closest(sc, orc, exact, "maxT == t")
    below = exact.copy()
    below[:, 7] = np.nextafter(t[hit], np.float32(0))
    _check_closest(sc, orc, below, "maxT < t")
    _check_any(sc, orc, below, "maxT < t any")
    prim, t0 = sc.trace(np.zeros((0, 8), np.float32))
    assert len(prim) == 0


@pytest.mark.parametrize("offset", [1.0e4, 1.0e5])
def test_trace_far_from_origin(ctx, tmp_path, offset):
    """The sphere scene translated to (offset, offset, offset), rays starting
    there: the slab test's rounding must scale with the distances near the
    boxes, not with the coordinates (at 1e5 the float spacing is 2^-7, far above
    the primitives' 1e-4 padding, which then rounds away)."""
    from cadise_amd.api import Scene, bvh_opts
    spec = small_spheres(n_spheres=8, level=3, offset=(offset, offset, offset))
    orc = OracleScene(spec)
    lo, hi = (offset - 18, offset, offset - 18), (offset + 18, offset + 8, offset + 18)
    rays = np.concatenate([random_rays(60000, lo, hi, seed=15), camera_rays(spec, 20000, seed=16)])
    # rays whose closest hit the reference's own BVH misses or changes (its slab
    # test culls a box holding the hit once the 1e-4 padding rounds away): listed
    # and counted separately; there the GPU must give the brute-force hit or the
    # reference BVH's
    bp, bt = orc.trace(rays, use_bvh=False)
    op, ot = orc.trace(rays)
    culled = (bp != op) | (bt.view(np.int32) != ot.view(np.int32))
    print(f"[offset {offset:g}] reference-BVH culls: {int(culled.sum())} of {len(rays)} rays "
          f"{np.flatnonzero(culled)[:8].tolist()}", flush=True)
    assert culled.sum() <= 1e-4 * len(rays)
    for width in (4, 8):
        sc = Scene.from_crsd(ctx, _write(tmp_path, spec), bvh_opts(width=width))
        _check_closest(sc, orc, rays[~culled], f"offset {offset:g} bvh{width}")
        gp, gt = sc.trace(rays[culled])
        ok = ((gp == bp[culled]) & (gt.view(np.int32) == bt[culled].view(np.int32))) | \
             ((gp == op[culled]) & (gt.view(np.int32) == ot[culled].view(np.int32)))
        print(f"[offset {offset:g} bvh{width}] culled rays: gpu {gp.tolist()} brute {bp[culled].tolist()} "
              f"reference bvh {op[culled].tolist()}", flush=True)
        assert ok.all()
        short = rays.copy()
        short[:, 7] = np.random.default_rng(17).uniform(0.1, 20.0, len(rays)).astype(np.float32)
        ab, _ = orc.trace(short, any_hit=True, use_bvh=False)
        ao, _ = orc.trace(short, any_hit=True)
        assert (ab != ao).sum() <= 1e-4 * len(rays)
        _check_any(sc, orc, short[ab == ao], f"offset {offset:g} bvh{width} any")


def _write(tmp_path, spec):
    path = tmp_path / "scene.crsd"
    path.write_text(spec.to_crsd(sidecar_dir=str(tmp_path)))
    return str(path)


def test_trace_degenerate_rays(ctx):
    """Rays the reference's tests never make: tmin > tmax, negative and zero tmax
    (an absent ray on the GPU returns no hit and t = tmax, as the reference's
    untouched maxT), infinite tmax, an infinite origin, tmin = -inf (hits behind
    the origin count, as in the reference), a zero direction. Ids and t as the
    oracle's. Rays with NaN components must come back (no hang, ids in range),
    but their ids are not compared: with NaN every comparison of the reference's
    tests passes, so its answer is whichever primitive it tests last, an
    artefact of traversal order."""
    from cadise_amd.api import Scene
    spec = scenes.cornell_box()
    sc = Scene.from_crsd_text(ctx, spec.to_crsd())
    orc = OracleScene(spec)
    F = np.finfo(np.float32).max
    rays = np.array([[0.5, 0.5, 0.5, 2.0, 0, 1, 0, 1.0], [0.5, 0.5, 0.5, 1e-3, 0, 1, 0, -1.0],
                     [0.5, 0.5, 0.5, 1e-3, 0, -1, 0, 0.0], [0.5, 0.5, 0.5, 1e-3, 0, 1, 0, np.inf],
                     [np.inf, 0.5, 0.5, 1e-3, -1, 0, 0, F], [0.5, 0.5, 0.5, -np.inf, 0, 1, 0, F],
                     [0.5, 0.5, 0.5, 1e-3, 0, 0, 0, F], [0.3, 0.2, 0.9, 0.0, 0, 0, -1, F]], np.float32)
    for any_hit in (False, True):
        prim, t = sc.trace(rays, any_hit)
        oprim, ot = orc.trace(rays, any_hit=any_hit)
        assert np.array_equal(prim, oprim), (any_hit, prim, oprim)
        if not any_hit:
            assert np.array_equal(t.view(np.int32), ot.view(np.int32)), (t, ot)
    rng = np.random.default_rng(31)
    mix = random_rays(20000, -0.1, 1.1, seed=32)
    bad = rng.random((len(mix), 8)) < 0.05
    mix[bad] = rng.choice(np.array([np.nan, np.inf, -np.inf, 0.0], np.float32), int(bad.sum()))
    for any_hit in (False, True):
        prim, _ = sc.trace(mix, any_hit)
        assert np.all((prim >= -1) & (prim < spec.num_prims())) if not any_hit else np.all((prim == 0) | (prim == 1))


def _golden_names():
    from tests.golden.make_golden import scene_set
    return [n for n, _ in scene_set()]


@pytest.mark.parametrize("name", _golden_names())
def test_surface_detail_vs_reference_golden(ctx, name):
    """Scene::isIntersecting's surface intersection (scene.cpp:30-48;
    triangle.cpp:125-160, rectangle.cpp:102-132, sphere.cpp:109-130) through
    mpt_intersect_detail, against the vectors the compiled reference wrote for
    the golden rays (tests/golden/<name>.npz "detail": primitive, t, P, Ng, Ns,
    uvw). Bit-identical for triangles and rectangles; a sphere's uvw comes from
    acos / atan2 of its normal, whose last bits differ between the device's and
    glibc's functions (a few ulps of an angle up to pi, which v = (pi - theta) /
    pi turns into more ulps near 0), so it may differ by 5e-7 absolute (its P, Ng
    and Ns are bit-identical). Rays with an exact-t tie compare t only; a ray that leaves the
    scene is a miss on both sides (the reference reports its background sphere
    as primitive -2)."""
    import os
    from cadise_amd.api import Scene
    from cadise_amd.scenes import PRIM_SPHERE
    from tests.golden.make_golden import scene_set
    spec = dict(scene_set())[name]
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    rays, ref = g["rays"], g["detail"]
    out = Scene.from_spec(ctx, spec).intersect_detail(rays)
    ties = OracleScene(spec).ties(rays)
    bits = lambda a: np.ascontiguousarray(a, np.float32).view(np.int32)
    rp, gp = ref[:, 0].astype(np.int64), out[:, 0].astype(np.int64)
    hit = rp >= 0
    assert np.array_equal(gp >= 0, hit), f"{name}: {int(((gp >= 0) != hit).sum())} hit / miss mismatches"
    assert np.all(out[~hit] == np.concatenate([[-1.0], np.zeros(13)]).astype(np.float32))
    assert np.array_equal(bits(out[hit, 1]), bits(ref[hit, 1])), f"{name}: t not bit-identical"
    uniq = hit & (ties <= 1)
    assert np.array_equal(gp[uniq], rp[uniq]), f"{name}: primitive mismatches"
    sphere = np.zeros(len(rays), bool)
    sphere[uniq] = spec.prim_arrays()[0][rp[uniq]] == PRIM_SPHERE
    flat, sph = uniq & ~sphere, uniq & sphere
    assert np.array_equal(bits(out[flat, 2:14]), bits(ref[flat, 2:14])), \
        f"{name}: {int((bits(out[flat, 2:14]) != bits(ref[flat, 2:14])).any(axis=1).sum())} triangle / rectangle rows differ"
    assert np.array_equal(bits(out[sph, 2:11]), bits(ref[sph, 2:11])), f"{name}: sphere P / Ng / Ns differ"
    ulps = np.abs(bits(out[sph, 11:14]).astype(np.int64) - bits(ref[sph, 11:14]).astype(np.int64))
    dev = np.abs(out[sph, 11:14].astype(np.float64) - ref[sph, 11:14])
    print(f"[detail {name}] rays {len(rays)}: hits {int(hit.sum())}, tied {int((hit & (ties > 1)).sum())}, "
          f"triangle / rectangle {int(flat.sum())} bit-identical, sphere {int(sph.sum())} (uvw within "
          f"{dev.max(initial=0.0):.2e}, {int(ulps.max(initial=0))} ulps; {int((ulps > 0).any(axis=1).sum())} rows "
          f"not bit-identical)", flush=True)
    assert dev.max(initial=0.0) <= 5e-7, f"{name}: sphere uvw {dev.max():.2e} apart"
