"""The product library without a GPU: it loads, exports every entry point of
include/mpt.h, and its host half (CRSD reader, the reference's BVH algorithm,
camera matrices, area-light radiance) matches the compiled reference's golden
vectors and the oracle bit for bit. Error paths return codes, never abort.
Uses a host-only context (mpt_create(-1)): no device call is made.
"""
import os
import re

import numpy as np
import pytest

from cadise_amd import native, scenes
from cadise_amd.api import Context, Scene
from cadise_amd.native import MptError
from oracle.binding import OracleScene
from tests.golden.make_golden import scene_set

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.fixture(scope="module")
def ctx():
    return Context(-1)


def test_exports_every_declared_function():
    header = open(os.path.join(ROOT, "include", "mpt.h")).read()
    declared = set(re.findall(r"\b(mpt_[a-z0-9_]+)\s*\(", header))
    lib = native.load_library()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(native.EXPORTS) == declared


@pytest.mark.parametrize("name", [n for n, _ in scene_set()])
def test_reference_tree_matches_golden(ctx, name, tmp_path):
    spec = dict(scene_set())[name]
    sc = Scene.from_crsd_text(ctx, spec.to_crsd(sidecar_dir=str(tmp_path)))
    b = sc.binary_bvh()
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    assert np.array_equal(bits(b["bounds"]), bits(g["bvh_bounds"]))
    for k in ("leaf", "a", "b", "ordered"):
        assert np.array_equal(b[k], g["bvh_" + k]), k


@pytest.mark.parametrize("name", [n for n, _ in scene_set()])
def test_camera_and_lights_match_oracle(ctx, name, tmp_path):
    spec = dict(scene_set())[name]
    sc = Scene.from_crsd_text(ctx, spec.to_crsd(sidecar_dir=str(tmp_path)))
    O = OracleScene(spec)
    f2c, c2w = sc.camera_matrices()
    of2c, oc2w = O.camera_matrices()
    assert np.array_equal(bits(f2c), bits(of2c)) and np.array_equal(bits(c2w), bits(oc2w))
    assert np.array_equal(bits(sc.light_radiance()), bits(O.light_radiance()))


def test_arrays_and_crsd_build_the_same_scene(ctx, tmp_path):
    spec = scenes.procedural_spheres(n_spheres=4, level=2, width=16, height=16, spp=1, seed=3)
    a = Scene.from_spec(ctx, spec).stats()
    b = Scene.from_crsd_text(ctx, spec.to_crsd(sidecar_dir=str(tmp_path))).stats()
    for k in ("nprims", "nlights", "binary_nodes", "wide_nodes", "binary_depth", "wide_depth"):
        assert a[k] == b[k], k


def test_gpu_tree_is_a_valid_sah_tree(ctx):
    spec = scenes.procedural_spheres(n_spheres=6, level=3, width=16, height=16, spp=1, seed=11)
    st = Scene.from_spec(ctx, spec).stats()
    n = st["nprims"]
    assert st["binary_nodes"] <= 2 * n - 1
    assert st["wide_nodes"] >= 1 and st["wide_depth"] <= st["binary_depth"]


BASE = """film { [image-width int32 "8"] [image-height int32 "8"] [filter string "box"] }
camera { [type string "perspective-pinhole"] [position vector3r "0 0 5"] [direction vector3r "0 0 -1"]
         [up vector3r "0 1 0"] [fov real "40"] }
renderer { [type string "%s"] [sampler string "stratified"] [sample-number int32 "4"] [estimator string "%s"] }
accelerator { [type string "bvh"] }
light-cluster { [type string "uniform"] }
material { [type string "%s"] [name string "m"] [albedo rgb "0.5 0.5 0.5"] }
primitive { [type string "%s"] [name string "p"] [bsdf material "m"] [v1 vector3r "-1 -1 0"] [v2 vector3r "1 -1 0"]
            [v3 vector3r "0 1 0"] [center vector3r "0 0 0"] [radius real "1"] }
light { [type string "area"] [primitive primitive "p"] [color rgb "1 1 1"] [watt real "10"] }
"""


@pytest.mark.parametrize("renderer,estimator,material,prim,code", [
    ("sampling", "pt", "matte-lambertian", "triangle", 0),
    ("bdpt", "pt", "matte-lambertian", "triangle", -3),
    ("vpm", "pt", "matte-lambertian", "triangle", -3),
    ("sampling", "vpt", "matte-lambertian", "triangle", 0),
    ("sampling", "wdl", "matte-lambertian", "triangle", -3),
    ("sampling", "pt", "plastic", "triangle", 0),
    ("sampling", "pt", "glass", "rectangle", 0),
    ("sampling", "pt", "conductor-dielectric", "triangle", -3),
    # a sphere light: the reference cannot sample a sphere's surface (sphere.cpp:132-136)
    ("sampling", "pt", "matte-lambertian", "sphere", -3),
])
def test_crsd_support_and_error_codes(ctx, renderer, estimator, material, prim, code):
    text = BASE % (renderer, estimator, material, prim)
    if code == 0:
        sc = Scene.from_crsd_text(ctx, text)
        assert sc.stats()["nprims"] == 1
    else:
        with pytest.raises(MptError) as e:
            Scene.from_crsd_text(ctx, text)
        assert e.value.code == code


def test_crsd_malformed(ctx):
    with pytest.raises(MptError) as e:
        Scene.from_crsd_text(ctx, "film { [image-width int32 \"8\"] }")
    assert e.value.code == -2
    with pytest.raises(MptError) as e:
        Scene.from_crsd(ctx, "/nonexistent/scene.crsd")
    assert e.value.code == -5


def test_host_only_context_refuses_device_calls(ctx):
    sc = Scene.from_crsd_text(ctx, BASE % ("sampling", "pt", "matte-lambertian", "triangle"))
    with pytest.raises(MptError) as e:
        sc.render()
    assert e.value.code == -4
    with pytest.raises(MptError):
        sc.trace(np.zeros((4, 8), np.float32))


def test_reference_tree_matches_oracle_on_stress_instances(ctx):
    """The reference's SAH tree at 300k triangles of the traversal-stress
    geometry (six rotated displaced grids): topology, bounds and primitive
    order equal the oracle's (bvhBuilder.cpp:18-358)."""
    spec = scenes.traversal_stress(n_instances=6)
    b = Scene.from_spec(ctx, spec).binary_bvh()
    o = OracleScene(spec).bvh()
    assert np.array_equal(bits(b["bounds"]), bits(o["bounds"]))
    for k in ("leaf", "a", "b", "ordered"):
        assert np.array_equal(b[k], o[k]), k


def test_texture_layout_and_sharing(ctx):
    """Materials that share a texture share one parameter block; offsets tile the
    flat texture array in creation order."""
    spec = scenes.textured_spheres(n_spheres=4, level=1, tex_res=8, width=8, height=8, spp=1)
    sc = Scene.from_spec(ctx, spec)
    st = sc.stats()
    assert st["ntextures"] == 2 and set(sc.tex_index) == {"albedo", "rough"}
    total = 0
    for i in range(st["ntextures"]):
        w, h, c, o = sc.texture_info(i)
        assert o == total and (w, h) == (8, 8)
        total += w * h * c
    assert total == st["tex_elements"] == 8 * 8 * 4
    sc.set_texture(sc.tex_index["albedo"], np.full((8, 8, 3), 0.5, np.float32))
    with pytest.raises(MptError) as e:
        sc.adjoint(np.zeros((8, 8, 3), np.float32))
    assert e.value.code == -4


@pytest.mark.parametrize("name", ["sky", "sky_envfirst"])
def test_environment_light_host(ctx, name, tmp_path):
    """N3 host half: the .hdr reader plus the loader's flips give the oracle's
    texture bit for bit, and the light cluster's pick probabilities (environment
    flux with the scene-bound or the default radius) equal the compiled
    reference's."""
    spec = dict(scene_set())[name]
    sc = Scene.from_crsd_text(ctx, spec.to_crsd(sidecar_dir=str(tmp_path)))
    st = sc.stats()
    h, w = spec.env.shape[:2]
    env = [i for i in range(st["ntextures"]) if sc.texture_info(i)[:3] == (w, h, 3)]
    assert len(env) == 1
    assert np.array_equal(bits(sc.get_texture(env[0])), bits(spec.env))
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    assert np.array_equal(bits(sc.light_pick_pdf()), bits(g["light_flux"][:, 1]))
    with pytest.raises(MptError) as e:
        sc.set_texture(env[0], spec.env)
    assert e.value.code == -3
    li = spec.light_order().index(("environment", None, -1))
    with pytest.raises(MptError) as e:
        sc.set_light_radiance(li, (1.0, 1.0, 1.0))
    assert e.value.code == -3


def test_environment_light_crsd_errors(ctx, tmp_path):
    spec = scenes.sky_spheres(n_spheres=1, level=0, width=8, height=8, spp=1, env_size=(16, 8))
    text = spec.to_crsd(sidecar_dir=str(tmp_path))
    hdr = os.path.join(str(tmp_path), "env.hdr")
    with pytest.raises(MptError) as e:       # only Radiance .hdr images
        Scene.from_crsd_text(ctx, text.replace(hdr, os.path.join(str(tmp_path), "env.png")))
    assert e.value.code == -3
    with pytest.raises(MptError) as e:       # missing image
        Scene.from_crsd_text(ctx, text.replace(hdr, os.path.join(str(tmp_path), "none.hdr")))
    assert e.value.code == -5
    block = f'light {{\n    [type string "environment"]\n    [hdr-filename string "{hdr}"]\n}}\n'
    with pytest.raises(MptError) as e:       # one background sphere
        Scene.from_crsd_text(ctx, text + block)
    assert e.value.code == -3
    with open(hdr, "r+b") as f:              # truncated pixel data
        data = f.read()
        f.seek(0)
        f.truncate()
        f.write(data[: len(data) // 2])
    with pytest.raises(MptError) as e:
        Scene.from_crsd_text(ctx, text)
    assert e.value.code in (-2, -5)


def test_wide_collapse_keeps_oversized_binary_leaves(ctx):
    """Three coincident triangles make one binary leaf of 3 (the reference's SAH
    cannot split them), above the default leaf size of 2: the collapse keeps that
    leaf whole and still builds a real tree over the other 70,000 triangles (it
    once turned every ancestor of such a leaf into one leaf, which no node can
    encode beyond 65535 primitives)."""
    rng = np.random.default_rng(3)
    c = rng.uniform(0, 100, (70000, 1, 3)).astype(np.float32)
    tris = c + rng.normal(0, 0.2, (70000, 3, 3)).astype(np.float32)
    dup = np.repeat(np.array([[[50, 50, 50], [51, 50, 50], [50, 51, 50]]], np.float32), 3, 0)
    spec = scenes.SceneSpec(width=8, height=8, spp=1, cam_pos=(50, 50, 200), cam_dir=(0, 0, -1), cam_up=(0, 1, 0))
    spec.materials = [scenes.Material("m")]
    spec.meshes = [scenes.Mesh("soup", "m", v=np.concatenate([tris, dup]).reshape(-1, 9)),
                   scenes.Mesh("lamp", "m", v=scenes.quad_tris((40, 120, 40), (60, 120, 40), (60, 120, 60), (40, 120, 60)))]
    spec.lights = [scenes.Light("lamp", (1, 1, 1), 100.0)]
    from cadise_amd.api import bvh_opts
    for width in (4, 8):
        for leaf_max in (1, 2):
            st = Scene.from_spec(ctx, spec, bvh_opts(width=width, leaf_max=leaf_max)).stats()
            assert st["wide_nodes"] > 70000 // (2 * width), (width, leaf_max, st["wide_nodes"])


def test_trace_ray_count_cap(ctx):
    """mpt_trace / mpt_trace_device refuse more than MPT_TRACE_MAX_RAYS = 2^31 - 2^24
    rays (the traversal's 32-bit work counter stays below 2^31 even when every
    wave over-reserves once) before any device work; at the cap the host-only
    context gets as far as its device check. The buffers are never read."""
    import ctypes as C
    header = open(os.path.join(ROOT, "include", "mpt.h")).read()
    assert "#define MPT_TRACE_MAX_RAYS ((size_t)2147483648u - (size_t)16777216u)" in header
    cap = (1 << 31) - (1 << 24)
    sc = Scene.from_spec(ctx, scenes.cornell_box(width=8, height=8, spp=1))
    L = native.load_library()
    dummy = np.zeros(16, np.float32)
    for fn, args in ((L.mpt_trace, lambda n: (sc.h, dummy.ctypes.data, n, 0, dummy.ctypes.data)),
                     (L.mpt_trace_device, lambda n: (sc.h, dummy.ctypes.data, n, 0, dummy.ctypes.data, None))):
        assert fn(*args(cap + 1)) == native.MPT_ERR_INVALID
        assert fn(*args(1 << 40)) == native.MPT_ERR_INVALID
        assert fn(*args(cap)) == native.MPT_ERR_DEVICE
    assert L.mpt_trace(sc.h, None, 5, 0, None) == native.MPT_ERR_INVALID


def test_primitive_count_cap(ctx):
    """Scenes of more than MPT_MAX_PRIMITIVES primitives are refused before any
    array is read or device work is done: the device addresses a primitive's
    three 16-byte records with 32-bit index arithmetic (3 * index + 2 < 2^31).
    The description's arrays hold one primitive; only the count says more."""
    import ctypes as C
    header = open(os.path.join(ROOT, "include", "mpt.h")).read()
    cap = int(re.search(r"#define MPT_MAX_PRIMITIVES \(\(int64_t\)(\d+)\)", header).group(1))
    assert 3 * (cap - 1) + 2 < (1 << 31) <= 3 * cap + 2
    t = np.zeros(1, np.int32)
    v = np.zeros(9, np.float32)
    m = np.zeros(1, np.int32)
    mats = (native.MaterialDesc * 1)()
    lights = (native.LightDesc * 1)()
    lights[0].prim = 0
    lights[0].color[:] = [1.0, 1.0, 1.0]
    lights[0].watt = 1.0
    d = native.SceneDesc()
    d.prim_type, d.prim_v, d.prim_mat = t.ctypes.data, v.ctypes.data, m.ctypes.data
    d.nmats, d.mats = 1, mats
    d.nlights, d.lights = 1, lights
    d.cam_dir[:] = [0.0, 0.0, -1.0]
    d.cam_up[:] = [0.0, 1.0, 0.0]
    d.fov_deg, d.sensor_width_mm = 40.0, 36.0
    d.width = d.height = d.tile_x = d.tile_y = 8
    d.spp, d.max_depth = 1, 4
    d.output_filename = b"x.png"
    L = native.load_library()
    h = C.c_void_p()
    for n in (cap + 1, 1 << 40):
        d.nprims = n
        assert L.mpt_scene_create(ctx.h, C.byref(d), None, C.byref(h)) == native.MPT_ERR_UNSUPPORTED
        assert "MPT_MAX_PRIMITIVES" in L.mpt_last_error(ctx.h).decode()
