"""The oracle against the golden vectors of the compiled reference
(tests/golden/make_golden.py). Bar: bit-exact (integer ids, float bit
patterns), on every fixture:

  A1/A2  BVH topology: node bounds, leaf flags, child / primitive indices,
         ordered primitive list (SAH and equal split)
  A3-A6  closest hit (primitive, t) and any hit on seeded rays
  A8     surface detail at the hit (P, Ng, Ns, uvw)
  A9     camera primary rays
  A11-A17 per-sample PtEstimator radiance and the number of uniforms drawn,
         with the same random stream
  A14-A16 BSDF evaluate / pdf / sample and area-light sampling on explicit
         uniforms; light emittance and pdf at hits
  A10/A18 the sampling renderer's linear film (stratified sampler, box and
         Gaussian filters, tile clamp)
"""
import os

import numpy as np
import pytest

from oracle.binding import OracleScene
from tests.golden.make_golden import scene_set

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCENES = dict(scene_set())


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.fixture(scope="module", params=sorted(SCENES))
def case(request):
    name = request.param
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    return name, OracleScene(SCENES[name]), g


def test_bvh_topology(case):
    name, O, g = case
    b = O.bvh()
    assert np.array_equal(bits(b["bounds"]), bits(g["bvh_bounds"])), name
    for k in ("leaf", "a", "b", "ordered"):
        assert np.array_equal(b[k], g["bvh_" + k]), (name, k)


def test_closest_and_any_hit(case):
    name, O, g = case
    prim, t = O.trace(g["rays"])
    assert np.array_equal(prim, g["closest_prim"]), name
    assert np.array_equal(bits(t), bits(g["closest_t"])), name
    short = g["rays"].copy()
    short[:, 7] = 2.0
    occ, _ = O.trace(short, any_hit=True)
    assert np.array_equal(occ, g["anyhit_maxT2"]), name


def test_surface_detail(case):
    name, O, g = case
    assert np.array_equal(bits(O.intersect_detail(g["rays"])), bits(g["detail"])), name


def test_camera_rays(case):
    name, O, g = case
    assert np.array_equal(bits(O.camera_rays(g["film_xy"])), bits(g["camera"])), name


def test_estimator_per_sample(case):
    name, O, g = case
    rgb, draws = O.estimate(g["film_xy"], g["est_keys"])
    assert np.array_equal(draws, g["est_draws"]), name
    assert np.array_equal(bits(rgb), bits(g["est_rgb"])), name


def test_render_linear_film(case):
    name, O, g = case
    film = O.render_linear(31, threads=4)
    assert np.array_equal(bits(film), bits(g["film"])), name


def test_light_hit(case):
    name, O, g = case
    assert np.array_equal(bits(O.light_hit(g["rays"])), bits(g["light_hit"])), name


def test_bsdf_vectors(case):
    name, O, g = case
    if "bsdf_in" not in g:
        pytest.skip("no BSDF vectors for this scene")
    for mi in range(len(g["bsdf_in"])):
        out = O.bsdf_query(mi, g["bsdf_in"][mi])
        assert np.array_equal(bits(out), bits(g["bsdf_out"][mi])), (name, mi)


def test_light_sample_vectors(case):
    name, O, g = case
    if "light_in" not in g:
        pytest.skip("no light vectors for this scene")
    for li in range(len(g["light_in"])):
        out = O.light_sample(li, g["light_in"][li], 16)
        assert np.array_equal(bits(out), bits(g["light_out"][li])), (name, li)
