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
  N3     environment light: background-sphere hits (uvw of the ray direction),
         emittance and pdf, Distribution2D sampling, the light's flux estimate
         with the scene-bound radius (declared last) or the default radius
         (declared first), and paths through the background
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


def test_light_flux_and_pick(case):
    """Light::approximateFlux of the environment light and the cluster's pick
    probabilities (power: flux-proportional; uniform: 1/n)."""
    name, O, g = case
    if "light_flux" not in g:
        pytest.skip("no environment light in this scene")
    flux, _, cdf = O.env_info()
    assert bits(flux) == bits(g["light_flux"][O.env_light, 0]), name
    pick = g["light_flux"][:, 1]
    if O.spec.cluster == "power":
        assert np.array_equal(bits(np.diff(O.power_cdf())), bits(pick)), name
    else:
        assert np.all(pick == np.float32(1.0) / np.float32(O.nlights)), name
    # the Distribution2D tables: nondecreasing rows ending at exactly 1
    h, w = O.spec.env.shape[:2]
    marg, cond = cdf[: h + 1], cdf[h + 1:].reshape(h, w + 1)
    assert marg[0] == 0 and marg[-1] == 1 and np.all(np.diff(marg) >= 0)
    assert np.all(cond[:, 0] == 0) and np.all(cond[:, -1] == 1) and np.all(np.diff(cond, axis=1) >= 0)
