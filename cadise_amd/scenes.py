"""Scene descriptions and the seeded workload generators.

A :class:`SceneSpec` is an in-memory CRSD scene (the reference's text scene
format, parsed by ``SdParser::parseSd``, source/file-io/scene-description/
sdParser.cpp:19-35). ``to_crsd`` writes it as CRSD text; large triangle meshes
go to raw float32 side-car files referenced by ``positions-file`` /
``normals-file`` / ``uvws-file`` keys of a ``triangle-mesh`` primitive, the
binary form this library's CRSD reader accepts beside the reference's
``vector3r-array`` text (primitiveInstantiator.cpp:64-90).

Generators (all seeded):
  * :func:`cornell_box`            36-triangle Cornell box, 0.25 x 0.25 light of radiance 17
  * :func:`procedural_spheres`     64 Perlin-displaced level-5 icospheres (1.31M triangles)
  * :func:`furnace_box`            closed Lambertian box whose walls all emit (closed form)
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

PRIM_TRIANGLE = 0
PRIM_RECTANGLE = 1
PRIM_SPHERE = 2          # rows of (cx, cy, cz, r, 0, 0, 0, 0, 0)

# CRSD material types of the path and their codes in the C ABI (MPT_MAT_*)
MAT_CODES = {"matte-lambertian": 0, "conductor-microfacet": 1, "phong": 2, "plastic": 3, "mirror": 4, "glass": 5,
             "transmission": 6, "absorb": 7}


def _f(x: float) -> str:
    # 9 significant digits round-trip a binary32 through std::stold + cast
    return "%.9g" % float(np.float32(x))


def _v(v) -> str:
    return " ".join(_f(x) for x in v)


@dataclass
class Material:
    name: str
    type: str = "matte-lambertian"        # a key of MAT_CODES
    albedo: tuple = (0.5, 0.5, 0.5)       # lambert albedo / plastic diffuse-albedo / specular albedo
    albedo_texture: str | None = None     # name of a texture-spectrum block
    f0: tuple = (1.0, 1.0, 1.0)           # schlick f0 (conductor, conductor mirror)
    roughness: float = 0.5
    roughness_texture: str | None = None  # name of a texture-real block
    exponent: float = 32.0                # phong exponent / plastic specular-exponent
    ratio: float = 0.7                    # plastic diffuse-ratio
    fresnel: str = "conductor"            # mirror: "conductor" (f0) or "dielectric" (iors)
    ior_outer: float = 1.0
    ior_inner: float | None = None        # default 1.5 (glass, dielectric mirror), 1.0 (transmission)

    def ior_inner_value(self) -> float:
        if self.ior_inner is not None:
            return float(self.ior_inner)
        return 1.0 if self.type == "transmission" else 1.5


@dataclass
class Texture:
    name: str
    data: np.ndarray | None               # H x W x C float32, row 0 = v in [0, 1/H) (None: checkerboard)
    kind: str = "spectrum"                # "spectrum" (3 channels) or "real" (1 channel)
    wrap: str = "repeat"
    file: str | None = None               # picture (.png / .tga / .hdr) written as an "image" texture;
                                          # data = its texels as the reference decodes them
    checker: tuple | None = None          # (odd-number, even-number, odd, even): odd / even a texture
                                          # name or a constant (rgb tuple, or a float for "real")


@dataclass
class Mesh:
    """A group of primitives sharing a name and a material."""
    name: str
    material: str
    kind: int = PRIM_TRIANGLE             # PRIM_TRIANGLE, PRIM_RECTANGLE or PRIM_SPHERE
    v: np.ndarray = None                  # n x 9 float32 (rectangle: v1 v2 v3)
    normals: np.ndarray | None = None     # n x 9 float32 or None
    uvws: np.ndarray | None = None        # n x 9 float32 or None
    as_mesh: bool = False                 # write as one triangle-mesh block


@dataclass
class Light:
    primitive: str                        # name of the emitting primitive
    color: tuple = (1.0, 1.0, 1.0)
    watt: float = 1.0
    back_face: bool = False


@dataclass
class SceneSpec:
    width: int = 256
    height: int = 256
    filter: str = "box"
    filter_width: float | None = None
    filter_height: float | None = None
    # other filter keys of the film block ("sigma-x", "sigma-y", "amplitude",
    # "b", "c"), absent = the reference's defaults (filterInstantiator.cpp:16-57)
    filter_params: dict = field(default_factory=dict)
    tile: int = 16
    cam_pos: tuple = (0.0, 0.0, 5.0)
    cam_dir: tuple = (0.0, 0.0, -1.0)
    cam_up: tuple = (0.0, 1.0, 0.0)
    fov: float = 40.0
    sampler: str = "stratified"
    spp: int = 16
    max_depth: int = 8
    split_mode: str = "sah"
    cluster: str = "uniform"
    output: str = "out.png"
    materials: list = field(default_factory=list)
    textures: list = field(default_factory=list)
    meshes: list = field(default_factory=list)
    lights: list = field(default_factory=list)
    # environment light: H x W x 3 radiance in texture orientation (row 0 = v in
    # [0, 1/H), the image as the reference's loader leaves it after its vertical
    # and horizontal flips), RGBE-exact (rgbe_quantize); declared before area
    # light env_slot (None: after all area lights)
    env: np.ndarray | None = None
    env_slot: int | None = None
    estimator: str = "pt"                 # "pt" (NEE + MIS) or "vpt" (BSDF sampling only)

    # -- derived views ------------------------------------------------------
    def num_prims(self) -> int:
        return int(sum(len(m.v) for m in self.meshes))

    def prim_arrays(self):
        """Creation-order primitive arrays as the CRSD reader produces them."""
        types, vs, ns, us, mats = [], [], [], [], []
        mat_index = {m.name: i for i, m in enumerate(self.materials)}
        for m in self.meshes:
            n = len(m.v)
            types.append(np.full(n, m.kind, np.int32))
            vs.append(np.asarray(m.v, np.float32).reshape(n, 9))
            ns.append(np.zeros((n, 9), np.float32) if m.normals is None else np.asarray(m.normals, np.float32))
            if m.uvws is None:
                u = np.zeros((n, 9), np.float32)
                u[:, 0] = np.nan
            else:
                u = np.asarray(m.uvws, np.float32)
            us.append(u)
            mats.append(np.full(n, mat_index[m.material], np.int32))
        cat = lambda a, shape: np.concatenate(a) if a else np.zeros(shape, np.float32)
        return (cat(types, (0,)).astype(np.int32), cat(vs, (0, 9)), cat(ns, (0, 9)), cat(us, (0, 9)),
                cat(mats, (0,)).astype(np.int32))

    def light_order(self):
        """The lights in declaration order: ("area", Light, primitive index) or
        ("environment", None, -1)."""
        area = [("area", l, p) for l, p in zip(self.lights, self.light_prims())]
        if self.env is None:
            return area
        slot = len(area) if self.env_slot is None else self.env_slot
        return area[:slot] + [("environment", None, -1)] + area[slot:]

    def light_prims(self):
        """Primitive index bound to each light: the first primitive carrying the
        name (std::map::insert keeps the first, primitiveInstantiator.cpp:119-125)."""
        first, base = {}, 0
        for m in self.meshes:
            first.setdefault(m.name, base)
            base += len(m.v)
        return [first[l.primitive] for l in self.lights]

    # -- CRSD ---------------------------------------------------------------
    def to_crsd(self, sidecar_dir: str | None = None, sidecar_threshold: int = 64) -> str:
        out = []
        flt = [f'[filter string "{self.filter}"]']
        if self.filter_width is not None:
            flt.append(f'[filter-width real "{_f(self.filter_width)}"]')
        if self.filter_height is not None:
            flt.append(f'[filter-height real "{_f(self.filter_height)}"]')
        for k, v in self.filter_params.items():
            flt.append(f'[{k} real "{_f(v)}"]')
        out.append("film {\n"
                   f'    [image-width int32 "{self.width}"]\n    [image-height int32 "{self.height}"]\n'
                   f'    [tile-size-x int32 "{self.tile}"]\n    [tile-size-y int32 "{self.tile}"]\n'
                   f'    [output-filename string "{self.output}"]\n    ' + "\n    ".join(flt) + "\n}")
        out.append("camera {\n    [type string \"perspective-pinhole\"]\n"
                   f'    [position vector3r "{_v(self.cam_pos)}"]\n    [direction vector3r "{_v(self.cam_dir)}"]\n'
                   f'    [up vector3r "{_v(self.cam_up)}"]\n    [fov real "{_f(self.fov)}"]\n}}')
        out.append("renderer {\n    [type string \"sampling\"]\n"
                   f'    [sampler string "{self.sampler}"]\n    [sample-number int32 "{self.spp}"]\n'
                   f'    [estimator string "{self.estimator}"]\n    [max-depth int32 "{self.max_depth}"]\n}}')
        out.append(f'accelerator {{\n    [type string "bvh"]\n    [split-mode string "{self.split_mode}"]\n}}')
        out.append(f'light-cluster {{\n    [type string "{self.cluster}"]\n}}')
        for t in self.textures:
            blk = "texture-spectrum" if t.kind == "spectrum" else "texture-real"
            if t.checker is not None:
                on, en, odd, even = t.checker

                def sub(key, x):
                    if isinstance(x, str):
                        return f'[{key} texture "{x}"]'
                    return f'[{key} rgb "{_v(x)}"]' if t.kind == "spectrum" else f'[{key} real "{_f(x)}"]'
                out.append(f'{blk} {{\n    [type string "checkerboard"]\n    [name string "{t.name}"]\n'
                           f'    [odd-number real "{_f(on)}"]\n    [even-number real "{_f(en)}"]\n'
                           f'    {sub("odd-texture", odd)}\n    {sub("even-texture", even)}\n}}')
                continue
            if t.file is not None:
                out.append(f'{blk} {{\n    [type string "image"]\n    [name string "{t.name}"]\n'
                           f'    [filename string "{os.path.abspath(t.file)}"]\n    [sample-mode string "nearest"]\n'
                           f'    [wrap-mode string "{t.wrap}"]\n}}')
                continue
            arr = np.ascontiguousarray(t.data, np.float32)
            if sidecar_dir is None:
                raise ValueError("image textures need a side-car directory")
            path = os.path.join(sidecar_dir, f"tex_{t.name}.f32")
            arr.tofile(path)
            out.append(f'{blk} {{\n    [type string "image-raw"]\n    [name string "{t.name}"]\n'
                       f'    [filename string "{path}"]\n    [width int32 "{arr.shape[1]}"]\n'
                       f'    [height int32 "{arr.shape[0]}"]\n    [wrap-mode string "{t.wrap}"]\n}}')
        for m in self.materials:
            head = f'material {{\n    [type string "{m.type}"]\n    [name string "{m.name}"]\n'

            def spectrum(key):
                return (f'[{key} texture "{m.albedo_texture}"]' if m.albedo_texture
                        else f'[{key} rgb "{_v(m.albedo)}"]')
            if m.type == "matte-lambertian":
                out.append(head + f'    {spectrum("albedo")}\n}}')
            elif m.type == "phong":
                out.append(head + f'    [exponent real "{_f(m.exponent)}"]\n}}')
            elif m.type == "plastic":
                out.append(head + f'    {spectrum("diffuse-albedo")}\n    [specular-exponent real "{_f(m.exponent)}"]\n'
                           f'    [diffuse-ratio real "{_f(m.ratio)}"]\n}}')
            elif m.type == "mirror":
                fr = (f'[fresnel-type string "conductor"]\n    [f0 rgb "{_v(m.f0)}"]' if m.fresnel == "conductor"
                      else f'[fresnel-type string "dielectric"]\n    [ior-outer real "{_f(m.ior_outer)}"]\n'
                           f'    [ior-inner real "{_f(m.ior_inner_value())}"]')
                out.append(head + f'    {spectrum("albedo")}\n    {fr}\n}}')
            elif m.type in ("glass", "transmission"):
                out.append(head + f'    {spectrum("albedo")}\n    [ior-outer real "{_f(m.ior_outer)}"]\n'
                           f'    [ior-inner real "{_f(m.ior_inner_value())}"]\n}}')
            elif m.type == "absorb":
                out.append(head + '}')
            else:
                rough = (f'[roughness texture "{m.roughness_texture}"]' if m.roughness_texture
                         else f'[roughness real "{_f(m.roughness)}"]')
                out.append(f'material {{\n    [type string "conductor-microfacet"]\n    [name string "{m.name}"]\n'
                           f'    [microfacet-type string "ggx"]\n    [fresnel-type string "schlick"]\n'
                           f'    [f0 rgb "{_v(m.f0)}"]\n    {rough}\n}}')
        for mi, m in enumerate(self.meshes):
            v = np.asarray(m.v, np.float32).reshape(-1, 9)
            if m.kind == PRIM_SPHERE:
                for r in v:
                    out.append(f'primitive {{\n    [type string "sphere"]\n    [name string "{m.name}"]\n'
                               f'    [bsdf material "{m.material}"]\n    [center vector3r "{_v(r[0:3])}"]\n'
                               f'    [radius real "{_f(r[3])}"]\n}}')
                continue
            if m.kind == PRIM_RECTANGLE or not m.as_mesh:
                kind = "rectangle" if m.kind == PRIM_RECTANGLE else "triangle"
                for r in v:
                    out.append(f'primitive {{\n    [type string "{kind}"]\n    [name string "{m.name}"]\n'
                               f'    [bsdf material "{m.material}"]\n    [v1 vector3r "{_v(r[0:3])}"]\n'
                               f'    [v2 vector3r "{_v(r[3:6])}"]\n    [v3 vector3r "{_v(r[6:9])}"]\n}}')
                continue
            lines = [f'primitive {{\n    [type string "triangle-mesh"]\n    [name string "{m.name}"]\n'
                     f'    [bsdf material "{m.material}"]']
            arrays = [("positions", v), ("normals", m.normals), ("uvws", m.uvws)]
            for key, arr in arrays:
                if arr is None:
                    continue
                arr = np.ascontiguousarray(np.asarray(arr, np.float32).reshape(-1, 9))
                if sidecar_dir is not None and len(arr) > sidecar_threshold:
                    path = os.path.join(sidecar_dir, f"mesh{mi}_{key}.f32")
                    arr.tofile(path)
                    lines.append(f'    [{key}-file string "{path}"]')
                else:
                    triples = ", ".join(_v(t) for t in arr.reshape(-1, 3))
                    lines.append(f'    [{key} vector3r-array "{triples}"]')
            out.append("\n".join(lines) + "\n}")
        for kind, l, _ in self.light_order():
            if kind == "environment":
                if sidecar_dir is None:
                    raise ValueError("an environment light needs a side-car directory")
                path = os.path.join(sidecar_dir, "env.hdr")
                write_hdr(path, self.env)
                out.append(f'light {{\n    [type string "environment"]\n    [hdr-filename string "{path}"]\n}}')
                continue
            out.append(f'light {{\n    [type string "area"]\n    [primitive primitive "{l.primitive}"]\n'
                       f'    [color rgb "{_v(l.color)}"]\n    [watt real "{_f(l.watt)}"]\n'
                       f'    [is-back-face-emit bool "{"true" if l.back_face else "false"}"]\n}}')
        return "# generated by cadise_amd.scenes\n" + "\n".join(out) + "\n"


# -- geometry helpers -----------------------------------------------------------

def quad_tris(a, b, c, d):
    """Two triangles (a, b, c), (a, c, d) of a planar quad a-b-c-d."""
    a, b, c, d = (np.asarray(x, np.float32) for x in (a, b, c, d))
    return np.stack([np.concatenate([a, b, c]), np.concatenate([a, c, d])]).astype(np.float32)


def box_tris(lo, hi):
    """12 triangles of an axis-aligned box, outward winding."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    p = lambda x, y, z: (x, y, z)
    faces = [
        (p(x0, y0, z0), p(x0, y0, z1), p(x0, y1, z1), p(x0, y1, z0)),  # -x
        (p(x1, y0, z0), p(x1, y1, z0), p(x1, y1, z1), p(x1, y0, z1)),  # +x
        (p(x0, y0, z0), p(x1, y0, z0), p(x1, y0, z1), p(x0, y0, z1)),  # -y
        (p(x0, y1, z0), p(x0, y1, z1), p(x1, y1, z1), p(x1, y1, z0)),  # +y
        (p(x0, y0, z0), p(x0, y1, z0), p(x1, y1, z0), p(x1, y0, z0)),  # -z
        (p(x0, y0, z1), p(x1, y0, z1), p(x1, y1, z1), p(x0, y1, z1)),  # +z
    ]
    return np.concatenate([quad_tris(*f) for f in faces])


def cornell_box(width=256, height=256, spp=16, max_depth=8, filter="box") -> SceneSpec:
    """Cornell box of 36 triangles: 5 walls, 2 axis-aligned boxes, and a
    0.25 x 0.25 ceiling light as two triangles, each its own area light of
    radiance 17 (watt = 3 * 17 * pi * area, singleAreaLight.cpp:20-37)."""
    s = SceneSpec(width=width, height=height, spp=spp, max_depth=max_depth, filter=filter,
                  cam_pos=(0.5, 0.5, 1.9), cam_dir=(0.0, 0.0, -1.0), cam_up=(0.0, 1.0, 0.0), fov=40.0)
    s.materials = [Material("white", albedo=(0.63, 0.63, 0.63)), Material("red", albedo=(0.63, 0.065, 0.05)),
                   Material("green", albedo=(0.14, 0.45, 0.091))]
    s.meshes = [
        Mesh("floor", "white", v=quad_tris((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0))),
        Mesh("ceiling", "white", v=quad_tris((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1))),
        Mesh("back", "white", v=quad_tris((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))),
        Mesh("left", "red", v=quad_tris((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1))),
        Mesh("right", "green", v=quad_tris((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0))),
        Mesh("shortbox", "white", v=box_tris((0.13, 0.0, 0.55), (0.43, 0.3, 0.85))),
        Mesh("tallbox", "white", v=box_tris((0.55, 0.0, 0.2), (0.85, 0.6, 0.5))),
    ]
    # light quad just below the ceiling, wound so that its normal points down
    y = 0.999
    lt = quad_tris((0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625))
    s.meshes.append(Mesh("light0", "white", v=lt[0:1]))
    s.meshes.append(Mesh("light1", "white", v=lt[1:2]))
    area = 0.5 * 0.25 * 0.25
    s.lights = [Light("light0", (1, 1, 1), 3 * 17 * math.pi * area), Light("light1", (1, 1, 1), 3 * 17 * math.pi * area)]
    return s


def checker_cornell(picture: str, texels: np.ndarray, width=32, height=32, spp=4, max_depth=6) -> SceneSpec:
    """The Cornell box with picture and checkerboard textures: the back wall's albedo
    is a checkerboard (5 x 3 cells over uv) of an "image" texture read from
    `picture` (texels: its texels as the reference decodes them, for the paths that
    take arrays) and a constant; the floor is a GGX conductor whose roughness is a
    checkerboard of two constants; the tall box's albedo is a checkerboard of a
    checkerboard and a constant."""
    s = cornell_box(width=width, height=height, spp=spp, max_depth=max_depth)
    s.textures = [Texture("pic", np.ascontiguousarray(texels, np.float32), file=picture),
                  Texture("back", None, checker=(5.0, 3.0, "pic", (0.2, 0.6, 0.3))),
                  Texture("rough", None, kind="real", checker=(4.0, 4.0, 0.15, 0.7)),
                  Texture("inner", None, checker=(2.0, 2.0, (0.9, 0.9, 0.2), "pic")),
                  Texture("nested", None, checker=(3.0, 1.5, "inner", (0.1, 0.1, 0.8)))]
    s.materials += [Material("backmat", albedo_texture="back"),
                    Material("floormat", type="conductor-microfacet", f0=(0.9, 0.7, 0.5), roughness_texture="rough"),
                    Material("boxmat", albedo_texture="nested")]
    swap = {"back": "backmat", "floor": "floormat", "tallbox": "boxmat"}
    for m in s.meshes:
        m.material = swap.get(m.name, m.material)
    return s


def shipped_cornell(width=640, height=480, spp=32, max_depth=10000) -> SceneSpec:
    """The reference's shipped scene (scene/cornell-box.crsd) as a SceneSpec: five
    rectangle walls (red, blue, white Lambertian), a plastic sphere (MixedBsdf of a
    Lambertian 0.7/0.56/0 and Phong exponent 16, diffuse ratio 0.8), a GGX sphere
    with roughness 0, one 190 x 165 ceiling rectangle light of 5e6 W; gaussian
    filter, 32 spp stratified, pt, uniform light cluster, max-depth default 10000."""
    s = SceneSpec(width=width, height=height, spp=spp, max_depth=max_depth, filter="gaussian",
                  cam_pos=(278.0, 273.0, 800.0), cam_dir=(0.0, 0.0, -1.0), cam_up=(0.0, 1.0, 0.0), fov=49.0,
                  output="cornell-box.png")
    s.materials = [Material("plane", albedo=(0.7, 0.7, 0.7)), Material("redPlane", albedo=(0.75, 0.3, 0.3)),
                   Material("bluePlane", albedo=(0.3, 0.3, 0.75)),
                   Material("microfacet", type="conductor-microfacet", f0=(1.0, 1.0, 1.0), roughness=0.0),
                   Material("plastic", type="plastic", albedo=(0.7, 0.56, 0.0), exponent=16.0, ratio=0.8)]
    rect = lambda *v: np.array([np.concatenate(v)], np.float32)
    s.meshes = [
        Mesh("leftWall", "redPlane", kind=PRIM_RECTANGLE, v=rect((-90.0, 0.0, -649.2), (-90.0, 0.0, 0.0), (-90.0, 548.8, 0.0))),
        Mesh("rightWall", "bluePlane", kind=PRIM_RECTANGLE,
             v=rect((642.8, 0.0, 0.0), (642.8, 0.0, -649.2), (642.8, 548.8, -649.2))),
        Mesh("topWall", "plane", kind=PRIM_RECTANGLE, v=rect((646.0, 548.8, 0.0), (-90.0, 548.8, 0.0), (-90.0, 548.8, -649.2))),
        Mesh("backWall", "plane", kind=PRIM_RECTANGLE,
             v=rect((-90.0, 548.8, -649.2), (646.0, 548.8, -649.2), (646.0, 0.0, -649.2))),
        Mesh("downWall", "plane", kind=PRIM_RECTANGLE, v=rect((642.8, 0.0, 0.0), (-90.0, 0.0, 0.0), (-90.0, 0.0, -649.2))),
        Mesh("leftSphere", "plastic", kind=PRIM_SPHERE, v=rect((96.0, 102.0, -361.0), (102.0, 0.0, 0.0), (0.0, 0.0, 0.0))),
        Mesh("rightSphere", "microfacet", kind=PRIM_SPHERE,
             v=rect((410.0, 111.0, -216.0), (111.0, 0.0, 0.0), (0.0, 0.0, 0.0))),
        Mesh("ceilingLight", "plane", kind=PRIM_RECTANGLE,
             v=rect((373.0, 538.7, -167.0), (373.0, 538.7, -332.0), (183.0, 538.7, -332.0))),
    ]
    s.lights = [Light("ceilingLight", (1.0, 1.0, 1.0), 5000000.0)]
    return s


def material_zoo(width=64, height=48, spp=4, max_depth=8, estimator="pt", env=None, cluster="uniform",
                 textured=False) -> SceneSpec:
    """Every material of the path in one box: a glass sphere, a dielectric
    mirror, a conductor mirror quad, a transmission sphere, a Phong sphere, a
    plastic sphere (its diffuse albedo an image texture when textured), a GGX
    sphere, an absorbing sphere, a sphere inside the glass one, Lambertian walls
    (open front) and two rectangle lights; optionally an environment image."""
    s = SceneSpec(width=width, height=height, spp=spp, max_depth=max_depth, filter="box", fov=55.0,
                  cam_pos=(5.0, 4.0, 16.0), cam_dir=(0.0, -0.12, -1.0), cam_up=(0.0, 1.0, 0.0),
                  estimator=estimator, cluster=cluster)
    if textured:
        tex = smooth_noise_texture(16, 3, 0.2, 0.9, 0x201)
        s.textures.append(Texture("zooalb", np.ascontiguousarray(tex, np.float32), kind="spectrum"))
    s.materials = [
        Material("wall", albedo=(0.65, 0.65, 0.6)), Material("red", albedo=(0.7, 0.15, 0.1)),
        Material("green", albedo=(0.15, 0.6, 0.2)),
        Material("glass", type="glass", albedo=(0.95, 0.95, 0.95), ior_outer=1.0, ior_inner=1.5),
        Material("dmirror", type="mirror", albedo=(0.9, 0.9, 0.9), fresnel="dielectric", ior_outer=1.0, ior_inner=1.33),
        Material("cmirror", type="mirror", albedo=(0.95, 0.9, 0.8), fresnel="conductor", f0=(0.9, 0.8, 0.6)),
        Material("clear", type="transmission", albedo=(0.9, 0.95, 0.9), ior_outer=1.0, ior_inner=1.2),
        Material("phong", type="phong", exponent=24.0),
        Material("plastic", type="plastic", albedo=(0.2, 0.5, 0.8), albedo_texture="zooalb" if textured else None,
                 exponent=40.0, ratio=0.6),
        Material("ggx", type="conductor-microfacet", f0=(0.8, 0.6, 0.3), roughness=0.3),
        Material("black", type="absorb"),
        Material("lamp", albedo=(0.8, 0.8, 0.8)),
    ]
    rect = lambda a, b, c: np.array([np.concatenate([a, b, c])], np.float32)
    sph = lambda c, r: np.array([[c[0], c[1], c[2], r, 0, 0, 0, 0, 0]], np.float32)
    W = 10.0
    s.meshes = [
        Mesh("floor", "wall", kind=PRIM_RECTANGLE, v=rect((W, 0, 0), (0, 0, 0), (0, 0, -W))),
        Mesh("ceiling", "wall", kind=PRIM_RECTANGLE, v=rect((0, W, 0), (W, W, 0), (W, W, -W))),
        Mesh("back", "wall", kind=PRIM_RECTANGLE, v=rect((0, W, -W), (W, W, -W), (W, 0, -W))),
        Mesh("left", "red", kind=PRIM_RECTANGLE, v=rect((0, 0, -W), (0, 0, 0), (0, W, 0))),
        Mesh("right", "green", kind=PRIM_RECTANGLE, v=rect((W, 0, 0), (W, 0, -W), (W, W, -W))),
        Mesh("cmirror", "cmirror", kind=PRIM_RECTANGLE, v=rect((9.9, 1.0, -2.0), (9.9, 1.0, -8.0), (9.9, 6.0, -8.0))),
        Mesh("glass", "glass", kind=PRIM_SPHERE, v=sph((2.5, 2.0, -3.0), 1.6)),
        Mesh("inner", "red", kind=PRIM_SPHERE, v=sph((2.5, 2.0, -3.0), 0.5)),
        Mesh("dmirror", "dmirror", kind=PRIM_SPHERE, v=sph((7.2, 1.4, -6.5), 1.4)),
        Mesh("clear", "clear", kind=PRIM_SPHERE, v=sph((6.0, 1.0, -2.0), 1.0)),
        Mesh("phong", "phong", kind=PRIM_SPHERE, v=sph((4.0, 1.2, -7.5), 1.2)),
        Mesh("plastic", "plastic", kind=PRIM_SPHERE, v=sph((1.6, 1.1, -7.6), 1.1)),
        Mesh("ggx", "ggx", kind=PRIM_SPHERE, v=sph((8.3, 4.5, -2.2), 0.9)),
        Mesh("black", "black", kind=PRIM_SPHERE, v=sph((5.0, 6.5, -5.0), 0.7)),
        # lamp0 faces down; lamp1's normal faces the left wall and it emits from its back
        Mesh("lamp0", "lamp", kind=PRIM_RECTANGLE, v=rect((6.5, 9.98, -6.5), (3.5, 9.98, -6.5), (3.5, 9.98, -3.5))),
        Mesh("lamp1", "lamp", kind=PRIM_RECTANGLE, v=rect((0.02, 5.5, -6.5), (0.02, 5.5, -8.0), (0.02, 7.0, -8.0))),
    ]
    s.lights = [Light("lamp0", (1.0, 0.95, 0.9), 600.0), Light("lamp1", (0.8, 0.9, 1.0), 150.0, back_face=True)]
    if env is not None:
        s.env = env
    return s


def furnace_box(albedo=0.5, radiance=1.0, width=32, height=32, spp=4, max_depth=4) -> SceneSpec:
    """Closed cube, Lambertian albedo a on every wall, every wall triangle an
    area light of radiance Le emitting inwards. E[L] = Le * sum_{k=0}^{D} a^k
    for max depth D (every vertex adds one power of a)."""
    s = SceneSpec(width=width, height=height, spp=spp, max_depth=max_depth, filter="box",
                  cam_pos=(0.0, 0.0, 0.0), cam_dir=(0.0, 0.0, -1.0), cam_up=(0.0, 1.0, 0.0), fov=60.0)
    s.materials = [Material("wall", albedo=(albedo, albedo, albedo))]
    tris = box_tris((-1, -1, -1), (1, 1, 1))
    # box_tris winds outwards; emission must face inwards -> back_face emit
    for i, t in enumerate(tris):
        s.meshes.append(Mesh(f"w{i}", "wall", v=t[None, :]))
        a = np.asarray(t, np.float64).reshape(3, 3)
        area = 0.5 * np.linalg.norm(np.cross(a[1] - a[0], a[2] - a[0]))
        s.lights.append(Light(f"w{i}", (1, 1, 1), 3 * radiance * math.pi * area, back_face=True))
    return s


# -- procedural spheres ------------------------------------------------------------

def icosphere(level: int):
    """Unit icosphere: (vertices V x 3 float64, faces F x 3 int)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = np.asarray(verts, np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.asarray(faces, np.int64)
    for _ in range(level):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        edges.sort(axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        mids = v[uniq[:, 0]] + v[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mids])
        inv = inv.reshape(3, -1) + base
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = inv[0], inv[1], inv[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def _perlin3(p: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """Perlin's improved noise (K. Perlin, "Improving Noise", SIGGRAPH 2002), vectorized."""
    pi = np.floor(p).astype(np.int64)
    pf = p - pi
    pi &= 255
    fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)
    u, v, w = fade(pf[:, 0]), fade(pf[:, 1]), fade(pf[:, 2])

    def grad(h, x, y, z):
        h = h & 15
        a = np.where(h < 8, x, y)
        b = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
        return np.where(h & 1, -a, a) + np.where(h & 2, -b, b)

    X, Y, Z = pi[:, 0], pi[:, 1], pi[:, 2]
    x, y, z = pf[:, 0], pf[:, 1], pf[:, 2]
    P = perm
    A = P[X] + Y; AA = P[A] + Z; AB = P[A + 1] + Z
    B = P[X + 1] + Y; BA = P[B] + Z; BB = P[B + 1] + Z
    lerp = lambda t, a, b: a + t * (b - a)
    return lerp(w,
                lerp(v, lerp(u, grad(P[AA], x, y, z), grad(P[BA], x - 1, y, z)),
                     lerp(u, grad(P[AB], x, y - 1, z), grad(P[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(P[AA + 1], x, y, z - 1), grad(P[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(P[AB + 1], x, y - 1, z - 1), grad(P[BB + 1], x - 1, y - 1, z - 1))))


def procedural_spheres(seed=0x20261203, n_spheres=64, level=5, width=1920, height=1080, spp=64, max_depth=8,
                       with_normals=True, offset=(64.0, 0.0, 64.0)) -> SceneSpec:
    """64 subdivided icospheres (20 * 4^level triangles each) with vertices
    displaced along the normal by Perlin noise of amplitude 0.1 r; radii
    log-uniform in [0.2, 2]; centres rejection-sampled without overlap above a
    two-triangle ground plane; GGX conductors with roughness U[0.05, 1] and
    f0 U[0.1, 0.9]^3; four quad (rectangle) lights.

    `offset` places the scene in the positive octant: the reference's empty
    bounding box has max = FLT_MIN (tAabb.ipp:13-17), so any node holding
    negative coordinates reaches the origin and its SAH stops splitting; with
    positive coordinates its BVH is a normal SAH tree and its CPU timing fair."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm, perm]).astype(np.int64)
    uv, uf = icosphere(level)
    ox, oy, oz = offset
    s = SceneSpec(width=width, height=height, spp=spp, max_depth=max_depth, filter="box", fov=60.0,
                  cam_pos=(ox + 0.0, oy + 9.0, oz + 30.0), cam_dir=(0.0, -0.28, -1.0), cam_up=(0.0, 1.0, 0.0))
    s.materials.append(Material("ground", albedo=(0.5, 0.5, 0.5)))
    s.materials.append(Material("lightmat", albedo=(0.8, 0.8, 0.8)))
    extent = 16.0
    centres, radii = [], []
    while len(centres) < n_spheres:
        r = float(np.exp(rng.uniform(math.log(0.2), math.log(2.0))))
        c = np.array([ox + rng.uniform(-extent, extent), oy + r * 1.1 + rng.uniform(0.0, 3.0),
                      oz + rng.uniform(-extent, extent)])
        if all(np.linalg.norm(c - c2) > (r + r2) * 1.15 for c2, r2 in zip(centres, radii)):
            centres.append(c)
            radii.append(r)
    for i, (c, r) in enumerate(zip(centres, radii)):
        n = _perlin3(uv * 2.5 + rng.uniform(0, 100, 3), perm)
        disp = uv * (r * (1.0 + 0.1 * n))[:, None]
        pos = disp + c
        tri = pos[uf].reshape(-1, 9).astype(np.float32)
        nrm = None
        if with_normals:
            fn = np.cross(pos[uf[:, 1]] - pos[uf[:, 0]], pos[uf[:, 2]] - pos[uf[:, 0]])
            vn = np.zeros_like(pos)
            for k in range(3):
                np.add.at(vn, uf[:, k], fn)
            vn /= np.linalg.norm(vn, axis=1, keepdims=True)
            nrm = vn[uf].reshape(-1, 9).astype(np.float32)
        mname = f"ggx{i}"
        s.materials.append(Material(mname, type="conductor-microfacet", f0=tuple(rng.uniform(0.1, 0.9, 3)),
                                    roughness=float(rng.uniform(0.05, 1.0))))
        s.meshes.append(Mesh(f"sphere{i}", mname, v=tri, normals=nrm, as_mesh=True))
    g = 60.0
    s.meshes.append(Mesh("ground", "ground", v=quad_tris((ox - g, oy, oz - g), (ox - g, oy, oz + g),
                                                         (ox + g, oy, oz + g), (ox + g, oy, oz - g))))
    lights = [((-10, 14, -10), 3.0), ((10, 14, -10), 3.0), ((-10, 14, 10), 3.0), ((10, 14, 10), 3.0)]
    for i, ((x, y, z), h) in enumerate(lights):
        x, y, z = x + ox, y + oy, z + oz
        # rectangle v1 v2 v3 -> vA vB vC; normal cross(vA - vB, vC - vB) points down
        rect = np.array([[x - h, y, z - h, x - h, y, z + h, x + h, y, z + h]], np.float32)
        s.meshes.append(Mesh(f"lamp{i}", "lightmat", kind=PRIM_RECTANGLE, v=rect))
        s.lights.append(Light(f"lamp{i}", (1.0, 0.95, 0.85), 4000.0))
    return s


def sky_spheres(seed=0x20261203 ^ 0x6, n_spheres=64, level=5, width=1920, height=1080, spp=16, max_depth=8,
                env_size=(2048, 1024), sun_diameter_deg=0.5, lamps=1, env_slot=None, offset=(64.0, 0.0, 64.0),
                cluster="power") -> SceneSpec:
    """procedural_spheres lit by a procedural sky (sky_texture: gradient, ground
    colour, sun disc) as the environment light plus `lamps` of its rectangle
    lights; the camera looks nearly level so the horizon and sky are in view.
    env_slot places the environment light among the lights (None: last)."""
    s = procedural_spheres(seed=seed, n_spheres=n_spheres, level=level, width=width, height=height, spp=spp,
                           max_depth=max_depth, offset=offset)
    s.cam_dir = (0.0, -0.1, -1.0)
    keep = {f"lamp{i}" for i in range(lamps)}
    s.meshes = [m for m in s.meshes if not m.name.startswith("lamp") or m.name in keep]
    s.lights = [l for l in s.lights if l.primitive in keep]
    s.env = sky_texture(env_size[0], env_size[1], sun_diameter_deg=sun_diameter_deg)
    s.env_slot = env_slot
    s.cluster = cluster
    return s


# -- traversal stress ----------------------------------------------------------------

def displaced_grid(nx: int, nz: int, size_x: float, size_z: float, amp: float, perm: np.ndarray,
                   shift) -> np.ndarray:
    """An nx x nz quad grid (2 nx nz triangles) in the xz plane centred on the
    origin, heights = amp * two octaves of Perlin noise; returns float64
    triangles (2 nx nz) x 3 x 3."""
    xs = np.linspace(-size_x / 2, size_x / 2, nx + 1)
    zs = np.linspace(-size_z / 2, size_z / 2, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    p = np.stack([X.ravel() * 0.35 + shift[0], Z.ravel() * 0.35 + shift[1], np.full(X.size, shift[2])], 1)
    y = amp * (_perlin3(p, perm) + 0.5 * _perlin3(p * 2.0 + 17.0, perm))
    verts = np.stack([X.ravel(), y, Z.ravel()], 1)
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    a = (i * (nz + 1) + j).ravel()
    b, c = a + 1, a + (nz + 1)
    d = c + 1
    faces = np.concatenate([np.stack([a, c, b], 1), np.stack([b, c, d], 1)])
    return verts[faces]


def _random_rotations(rng, n):
    """n uniformly distributed rotation matrices (normalized Gaussian quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)


def traversal_stress(seed=0x20261203 ^ 0x3, n_instances=400, n_bases=4, grid=(125, 200), size=(12.0, 19.2),
                     amp=1.0, width=1024, height=1024, spp=4, max_depth=4, light="sky") -> SceneSpec:
    """Traversal stress scene: n_instances copies (flattened to world space) of
    n_bases seeded displaced grids of 2 * 125 * 200 = 50,000 triangles each,
    randomly rotated, scaled by U[0.6, 1.4] and translated into [15, 85]^3 of a
    100^3 box (400 instances: 20,000,000 triangles). Every coordinate is
    positive (the reference's empty-box quirk, tAabb.ipp:13-17, is then inert).
    Lambertian grey; the camera looks into the box from z = -35. Lit by the
    procedural sky (sky_texture: 2048 x 1024, gradient + 0.5-degree sun) as the
    environment light, or (light="rect") by one rectangle light at the top."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm, perm]).astype(np.int64)
    bases = [displaced_grid(grid[0], grid[1], size[0], size[1], amp, perm, rng.uniform(0, 50, 3))
             for _ in range(n_bases)]
    s = SceneSpec(width=width, height=height, spp=spp, max_depth=max_depth, filter="box", fov=60.0,
                  cam_pos=(50.0, 50.0, -35.0), cam_dir=(0.0, 0.0, 1.0), cam_up=(0.0, 1.0, 0.0))
    s.materials.append(Material("grey", albedo=(0.6, 0.6, 0.6)))
    rots = _random_rotations(rng, n_instances)
    scales = rng.uniform(0.6, 1.4, n_instances)
    trans = rng.uniform(15.0, 85.0, (n_instances, 3))
    which = rng.integers(0, n_bases, n_instances)
    for k in range(n_instances):
        tri = bases[which[k]] @ (rots[k].T * scales[k]) + trans[k]
        s.meshes.append(Mesh(f"inst{k}", "grey", v=tri.reshape(-1, 9).astype(np.float32), as_mesh=True))
    if light == "sky":
        s.env = sky_texture()
        return s
    # one 30 x 30 rectangle light at the top of the box, facing down
    s.materials.append(Material("lightmat", albedo=(0.8, 0.8, 0.8)))
    rect = np.array([[35.0, 99.5, 35.0, 35.0, 99.5, 65.0, 65.0, 99.5, 65.0]], np.float32)
    s.meshes.append(Mesh("sun", "lightmat", kind=PRIM_RECTANGLE, v=rect))
    s.lights.append(Light("sun", (1.0, 0.95, 0.85), 200000.0))
    return s


def surface_rays(spec: SceneSpec, n: int, seed: int) -> np.ndarray:
    """n x 8 secondary rays: origins at uniform points of random triangles,
    directions cosine-distributed about the geometric normal (random side),
    tmin 1e-3, tmax FLT_MAX. Directions are unit length in float32 but not
    normalized the reference's way (callers that need that re-normalize)."""
    rng = np.random.default_rng(seed)
    counts = np.array([len(m.v) for m in spec.meshes])
    starts = np.concatenate([[0], np.cumsum(counts)])
    idx = rng.integers(0, starts[-1], n)
    mesh = np.searchsorted(starts, idx, side="right") - 1
    v = np.empty((n, 9), np.float64)
    for mi in np.unique(mesh):
        sel = mesh == mi
        v[sel] = spec.meshes[mi].v[idx[sel] - starts[mi]]
    a, b, c = v[:, 0:3], v[:, 3:6], v[:, 6:9]
    u1, u2 = rng.random(n), rng.random(n)
    su = np.sqrt(u1)
    p = a * (1 - su)[:, None] + b * (su * (1 - u2))[:, None] + c * (su * u2)[:, None]
    nrm = np.cross(b - a, c - a)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    nrm *= rng.choice([-1.0, 1.0], n)[:, None]
    # cosine-weighted direction in the frame of nrm
    r1, r2 = rng.random(n), rng.random(n)
    phi, sr = 2 * np.pi * r1, np.sqrt(r2)
    helper = np.where(np.abs(nrm[:, :1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    t1 = np.cross(helper, nrm)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(nrm, t1)
    d = t1 * (sr * np.cos(phi))[:, None] + t2 * (sr * np.sin(phi))[:, None] + nrm * np.sqrt(1 - r2)[:, None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty((n, 8), np.float32)
    rays[:, 0:3] = p
    rays[:, 3] = 1e-3
    rays[:, 4:7] = d
    rays[:, 7] = np.finfo(np.float32).max
    return rays


# -- differentiable-rendering scene ----------------------------------------------------

def smooth_noise_texture(res: int, channels: int, lo: float, hi: float, seed: int, freq: float = 4.0) -> np.ndarray:
    """res x res x channels float32 in [lo, hi]: two octaves of Perlin noise,
    tileable in neither direction (the texture is sampled nearest-texel)."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm, perm]).astype(np.int64)
    y, x = np.meshgrid(np.arange(res) + 0.5, np.arange(res) + 0.5, indexing="ij")
    out = np.empty((res, res, channels), np.float32)
    for c in range(channels):
        off = rng.uniform(0, 100, 3)
        p = np.stack([x.ravel() / res * freq + off[0], y.ravel() / res * freq + off[1], np.full(x.size, off[2])], 1)
        n = _perlin3(p, perm) + 0.5 * _perlin3(p * 2.0 + 31.0, perm)
        n = (n - n.min()) / max(n.max() - n.min(), 1e-9)
        out[..., c] = (lo + (hi - lo) * n).reshape(res, res)
    return out


def sphere_uvs(unit_verts: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """Per-corner (u, v, 0) of a spherical mapping, u = atan2(z, x) / 2pi + 1/2,
    v = acos(y) / pi; triangles straddling the u seam get u + 1 on their small
    corners (repeat wrap brings them back), so no triangle sweeps the texture."""
    u = np.arctan2(unit_verts[:, 2], unit_verts[:, 0]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(unit_verts[:, 1], -1, 1)) / np.pi
    fu = u[faces]
    seam = fu.max(1) - fu.min(1) > 0.5
    fu[seam] = np.where(fu[seam] < 0.5, fu[seam] + 1.0, fu[seam])
    uvw = np.zeros((len(faces), 3, 3), np.float64)
    uvw[:, :, 0] = fu
    uvw[:, :, 1] = v[faces]
    return uvw.reshape(-1, 9).astype(np.float32)


def textured_spheres(seed=0x20261203 ^ 0x5, n_spheres=10, level=5, tex_res=256, width=512, height=512, spp=16,
                     max_depth=4, albedo=None, roughness=None, light_radiance=None, light_half=2.0) -> SceneSpec:
    """Differentiation scene: n_spheres displaced icospheres (20 * 4^level
    triangles each; 10 x level 5 = 204,800) on a ground quad, spherical uvs;
    odd spheres Lambertian with the shared albedo texture (tex_res^2 RGB, smooth
    noise in [0.15, 0.85]), even spheres GGX conductors with the shared roughness
    texture (tex_res^2, smooth noise in [0.1, 0.9]); the ground uses the albedo
    texture too. One rectangle light whose emitted radiance is a parameter
    (watt chosen so Le = light_radiance, default (10, 10, 10)). Every coordinate is
    positive."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm, perm]).astype(np.int64)
    uv, uf = icosphere(level)
    o = np.array([20.0, 0.0, 20.0])
    s = SceneSpec(width=width, height=height, spp=spp, max_depth=max_depth, filter="box", fov=50.0,
                  cam_pos=tuple(o + [0.0, 4.5, 11.0]), cam_dir=(0.0, -0.38, -1.0), cam_up=(0.0, 1.0, 0.0))
    alb = smooth_noise_texture(tex_res, 3, 0.15, 0.85, seed + 1) if albedo is None else albedo
    rough = smooth_noise_texture(tex_res, 1, 0.1, 0.9, seed + 2) if roughness is None else roughness
    s.textures.append(Texture("albedo", np.ascontiguousarray(alb, np.float32), kind="spectrum"))
    s.textures.append(Texture("rough", np.ascontiguousarray(rough, np.float32).reshape(tex_res, tex_res, 1),
                              kind="real"))
    s.materials.append(Material("ground", albedo_texture="albedo"))
    s.materials.append(Material("lightmat", albedo=(0.8, 0.8, 0.8)))
    s.materials.append(Material("diffuse", albedo_texture="albedo"))
    centres, radii = [], []
    extent = 4.5
    while len(centres) < n_spheres:
        r = float(np.exp(rng.uniform(np.log(0.5), np.log(1.3))))
        c = np.array([o[0] + rng.uniform(-extent, extent), r * 1.05, o[2] + rng.uniform(-extent, extent * 0.6)])
        if all(np.linalg.norm(c - c2) > (r + r2) * 1.1 for c2, r2 in zip(centres, radii)):
            centres.append(c)
            radii.append(r)
    uvw = sphere_uvs(uv, uf)
    for i, (c, r) in enumerate(zip(centres, radii)):
        n = _perlin3(uv * 2.0 + rng.uniform(0, 100, 3), perm)
        pos = uv * (r * (1.0 + 0.05 * n))[:, None] + c
        tri = pos[uf].reshape(-1, 9).astype(np.float32)
        fn = np.cross(pos[uf[:, 1]] - pos[uf[:, 0]], pos[uf[:, 2]] - pos[uf[:, 0]])
        vn = np.zeros_like(pos)
        for k in range(3):
            np.add.at(vn, uf[:, k], fn)
        vn /= np.linalg.norm(vn, axis=1, keepdims=True)
        nrm = vn[uf].reshape(-1, 9).astype(np.float32)
        if i % 2:
            mname = "diffuse"
        else:
            mname = f"ggx{i}"
            s.materials.append(Material(mname, type="conductor-microfacet", f0=tuple(rng.uniform(0.3, 0.9, 3)),
                                        roughness_texture="rough"))
        s.meshes.append(Mesh(f"sphere{i}", mname, v=tri, normals=nrm, uvws=uvw, as_mesh=True))
    g = 14.0
    ground = quad_tris((o[0] - g, 0.0, o[2] - g), (o[0] - g, 0.0, o[2] + g), (o[0] + g, 0.0, o[2] + g),
                       (o[0] + g, 0.0, o[2] - g))
    guv = ((ground.reshape(-1, 3)[:, [0, 2]] - (o[[0, 2]] - g)) / (2 * g) * 2.0).reshape(-1, 3, 2)
    gu = np.zeros((len(ground), 3, 3), np.float32)
    gu[:, :, :2] = guv
    s.meshes.append(Mesh("ground", "ground", v=ground, uvws=gu.reshape(-1, 9)))
    # square rectangle light (side 2 * light_half), facing down, 7 above the ground
    h = light_half
    y = 7.0
    rect = np.array([[o[0] - h, y, o[2] - h, o[0] - h, y, o[2] + h, o[0] + h, y, o[2] + h]], np.float32)
    s.meshes.append(Mesh("lamp", "lightmat", kind=PRIM_RECTANGLE, v=rect))
    Le = (10.0, 10.0, 10.0) if light_radiance is None else tuple(light_radiance)
    # Le = color / sum(color) * watt / (area * pi) (singleAreaLight.cpp:32-36) with color = Le / sum(Le) * 3
    area = (2 * h) * (2 * h)
    tot = float(sum(Le))
    s.lights.append(Light("lamp", tuple(3.0 * x / tot for x in Le), tot * area * np.pi))
    return s


# -- batch of scenes -----------------------------------------------------------------------

def batch_scenes(seed=0x20261203 ^ 0x4, n_scenes=64, tris=(1e4, 2e6), res=(128, 1024), spp=(1, 64), depth=(4, 16)):
    """n_scenes procedural-sphere scenes of uneven size: triangle count and
    square resolution log-uniform in the given ranges, spp and max depth
    uniform integers, GGX roughness U[0.05, 1] (each scene as
    procedural_spheres with its own seed; the sphere level is the largest that
    keeps at most 100 spheres)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_scenes):
        t = float(np.exp(rng.uniform(np.log(tris[0]), np.log(tris[1]))))
        r = int(round(np.exp(rng.uniform(np.log(res[0]), np.log(res[1])))))
        s = int(rng.integers(spp[0], spp[1] + 1))
        d = int(rng.integers(depth[0], depth[1] + 1))
        level = 1
        for lv in (5, 4, 3, 2, 1):
            if t / (20 * 4 ** lv) >= 1:
                level = lv
                break
        while t / (20 * 4 ** level) > 100 and level < 6:
            level += 1
        n = int(max(1, min(100, round(t / (20 * 4 ** level)))))
        sp = procedural_spheres(seed=int(rng.integers(1, 2**31)), n_spheres=n, level=level, width=r, height=r, spp=s,
                                max_depth=d)
        out.append(sp)
    return out


# -- environment light images (Radiance RGBE) ------------------------------------------------

def rgbe_quantize(img: np.ndarray) -> np.ndarray:
    """The float32 values an RGBE file stores for img (H x W x 3): Ward's
    encoding (shared exponent of the largest channel, 8-bit mantissas) decoded
    as m * 2^(e - 136)."""
    x = np.asarray(img, np.float64)
    v = x.max(-1)
    m, e = np.frexp(v)
    ok = v >= 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    mant = np.floor(x * scale[..., None]).astype(np.int64).clip(0, 255)
    return (mant * np.ldexp(1.0, (e - 8))[..., None] * ok[..., None]).astype(np.float32)


def _rgbe_bytes(img: np.ndarray) -> np.ndarray:
    x = np.asarray(img, np.float64)
    v = x.max(-1)
    m, e = np.frexp(v)
    ok = v >= 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    out = np.zeros(x.shape[:-1] + (4,), np.uint8)
    out[..., :3] = np.floor(x * scale[..., None]).clip(0, 255).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    return out


def write_hdr(path: str, tex: np.ndarray) -> None:
    """Radiance .hdr (new-style run-length scanlines of literal runs) such that
    a loader that flips the image vertically (row 0 = bottom scanline) and then
    horizontally gets tex back (after rgbe_quantize)."""
    img = np.ascontiguousarray(np.asarray(tex, np.float32)[::-1, ::-1])   # file rows top to bottom
    h, w = img.shape[:2]
    px = _rgbe_bytes(img)
    chunks = [b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n", f"-Y {h} +X {w}\n".encode()]
    if w < 8 or w > 0x7FFF:
        chunks.append(px.tobytes())
    else:
        for row in px:
            buf = bytearray([2, 2, w >> 8, w & 255])
            for c in range(4):
                comp = row[:, c]
                for i in range(0, w, 128):
                    seg = comp[i:i + 128]
                    buf.append(len(seg))
                    buf.extend(seg.tobytes())
            chunks.append(bytes(buf))
    with open(path, "wb") as f:
        f.write(b"".join(chunks))


def sky_texture(width=2048, height=1024, sun_elevation_deg=40.0, sun_azimuth=0.3, sun_diameter_deg=0.5,
                sun_radiance=(2.0e4, 1.8e4, 1.5e4), zenith=(0.25, 0.45, 1.0), horizon=(1.0, 1.0, 1.1),
                ground=(0.25, 0.22, 0.2)) -> np.ndarray:
    """Procedural sky as the environment light's texture (RGBE-exact): a
    gradient from horizon to zenith colour over the upper hemisphere, a dim
    ground colour below, and a sun disc of the given angular diameter. Texel
    (x, y) is looked up at u = (x + 0.5) / W, v = (y + 0.5) / H, i.e. the
    direction with elevation (v - 1/2) pi and azimuth atan2(x, z) = 2 pi u (the
    environment light's lookup frame, y up)."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    elev = (v - 0.5) * np.pi
    azim = u * 2 * np.pi
    E, A = np.meshgrid(elev, azim, indexing="ij")
    up = np.clip(np.sin(E), 0.0, 1.0)[..., None]
    sky = np.asarray(horizon) * (1 - up) ** 3 + np.asarray(zenith) * (1 - (1 - up) ** 3)
    img = np.where(E[..., None] >= 0, sky, np.asarray(ground))
    d = np.stack([np.cos(E) * np.sin(A), np.sin(E), np.cos(E) * np.cos(A)], -1)
    se, sa = np.radians(sun_elevation_deg), sun_azimuth * 2 * np.pi
    s = np.array([np.cos(se) * np.sin(sa), np.sin(se), np.cos(se) * np.cos(sa)])
    ang = np.degrees(np.arccos(np.clip(d @ s, -1, 1)))
    img = np.where((ang <= sun_diameter_deg / 2)[..., None], np.asarray(sun_radiance), img)
    return rgbe_quantize(img.astype(np.float32))
