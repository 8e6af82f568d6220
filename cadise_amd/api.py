"""Python face of the C ABI, named after the reference's plugin points
(reference source/core/renderDatabase.h:25-28, core/renderer/renderer.h:17,
core/intersector/accelerator/accelerator.h:9-12):

  Scene.from_crsd / from_spec      RenderDatabase::setUpData + prepareRender
  Scene.is_intersecting / is_occluded   Accelerator::isIntersecting / isOccluded, batched
  Scene.render                      SamplingRenderer::render with the pt estimator
  Scene.save                        Film::save

Errors raise MptError (the reference aborts; this library returns codes).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .scenes import MAT_CODES
from .native import (MPT_RENDER_KEEP_SAMPLES, MPT_RENDER_PASSWISE, BatchOpts, BatchTrial, BvhOpts, LightDesc, MaterialDesc, MptError,
                     RenderParams, SceneDesc, SceneStats, Stats, Texture)

_FILTERS = {"box": 0, "gaussian": 1, "cone": 2, "mitchell": 3}
_FILTER_W = {"box": 1.0, "gaussian": 4.0, "cone": 4.0, "mitchell": 4.0}


class Context:
    """device >= 0: GPU context; device == -1: host-only (parse + BVH build, no upload)."""

    def __init__(self, device: int = 0):
        self.L = native.load_library()
        h = C.c_void_p()
        rc = self.L.mpt_create(device, C.byref(h))
        if rc != 0:
            raise MptError(rc, f"mpt_create(device={device}) failed (is a GPU visible?)")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise MptError(rc, self.L.mpt_last_error(self.h).decode())

    def release_stream_buffers(self):
        """Frees the streaming renderer's path slots and queues (reallocated on demand)."""
        self.check(self.L.mpt_release_stream_buffers(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.L.mpt_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def bvh_opts(width=4, leaf_max=2, block=0, stack_lds=0, blocks_per_cu=0):
    """Zero fields take the library defaults for the width (block 128; BVH4: 20 LDS stack entries,
    8 blocks per CU; BVH8: 16 entries, 7 blocks)."""
    return BvhOpts(width, leaf_max, block, stack_lds, blocks_per_cu)


class Scene:
    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.L = ctx.L
        self.h = handle
        st = SceneStats()
        ctx.check(self.L.mpt_scene_stats_get(self.h, C.byref(st)))
        self.width, self.height = st.image_width, st.image_height
        self.spp, self.max_depth = st.spp, st.max_depth
        self.tex_index = {}

    # -- construction -----------------------------------------------------------
    @classmethod
    def from_crsd_text(cls, ctx: Context, text: str, opts: BvhOpts | None = None):
        h = C.c_void_p()
        data = text.encode()
        ctx.check(ctx.L.mpt_scene_parse_crsd(ctx.h, data, len(data), C.byref(opts) if opts else None, C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_crsd(cls, ctx: Context, path: str, opts: BvhOpts | None = None):
        h = C.c_void_p()
        ctx.check(ctx.L.mpt_scene_load_crsd(ctx.h, path.encode(), C.byref(opts) if opts else None, C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_spec(cls, ctx: Context, spec, opts: BvhOpts | None = None):
        """Build from a cadise_amd.scenes.SceneSpec through mpt_scene_create (no text round trip)."""
        t, v, n, u, m = spec.prim_arrays()
        keep = [t, v, n, u, m]
        tex = {x.name: x for x in spec.textures}
        arrays = {}        # one array per texture name: materials sharing it share one device texture
        tex_index = {}     # image texture name -> index in the scene's texture list (creation order)
        created = [0]      # textures the library creates, in its order: a checkerboard's odd and
                           # even textures (constants included) before the checkerboard itself

        def texture(const, name, channels, depth=0):
            T = Texture()
            if name is None:
                c = const if isinstance(const, (tuple, list)) else (const, const, const)
                T.value[:] = [float(np.float32(x)) for x in c]
                created[0] += 1 if depth > 0 else 0
                return T
            if tex[name].checker is not None:
                on, en, odd, even = tex[name].checker
                kids = [texture(*sub_texture(x), channels, depth + 1) for x in (odd, even)]
                keep.extend(kids)
                T.odd, T.even = C.addressof(kids[0]), C.addressof(kids[1])
                T.odd_number, T.even_number = float(np.float32(on)), float(np.float32(en))
                created[0] += 1
                return T
            if name not in arrays:
                arrays[name] = np.ascontiguousarray(tex[name].data, np.float32)
                tex_index[name] = created[0]
                created[0] += 1
            arr = arrays[name]
            keep.append(arr)
            T.data = arr.ctypes.data
            T.height, T.width = arr.shape[0], arr.shape[1]
            T.channels = arr.shape[2] if arr.ndim == 3 else 1
            T.wrap_clamp = 1 if tex[name].wrap == "clamp" else 0
            return T

        def sub_texture(x):   # a checkerboard's odd / even: a texture name or a constant
            return (None, x) if isinstance(x, str) else (x, None)

        mats = (MaterialDesc * len(spec.materials))()
        for i, mat in enumerate(spec.materials):
            mats[i].type = MAT_CODES[mat.type]
            mats[i].albedo = texture(mat.albedo, mat.albedo_texture, 3)
            mats[i].f0[:] = [float(np.float32(x)) for x in mat.f0]
            mats[i].roughness = texture(mat.roughness, mat.roughness_texture, 1)
            mats[i].exponent = float(np.float32(mat.exponent))
            mats[i].ratio = float(np.float32(mat.ratio))
            mats[i].fresnel = 0 if mat.fresnel == "conductor" else 1
            mats[i].ior_outer = float(np.float32(mat.ior_outer))
            mats[i].ior_inner = float(np.float32(mat.ior_inner_value()))
        order = spec.light_order()
        lights = (LightDesc * max(1, len(order)))()
        for i, (kind, l, p) in enumerate(order):
            if kind == "environment":
                env = np.ascontiguousarray(spec.env, np.float32)
                keep.append(env)
                E = lights[i]
                E.type = 1
                E.prim = -1
                E.env.data = env.ctypes.data
                E.env.height, E.env.width, E.env.channels = env.shape[0], env.shape[1], 3
                continue
            lights[i].prim = p
            lights[i].color[:] = [float(np.float32(x)) for x in l.color]
            lights[i].watt = float(np.float32(l.watt))
            lights[i].back_face = int(l.back_face)
        d = SceneDesc()
        d.nprims = len(t)
        d.prim_type, d.prim_v, d.prim_mat = t.ctypes.data, v.ctypes.data, m.ctypes.data
        d.prim_n = n.ctypes.data if any(x.normals is not None for x in spec.meshes) else None
        d.prim_uvw = u.ctypes.data if any(x.uvws is not None for x in spec.meshes) else None
        d.nmats, d.mats = len(spec.materials), mats
        d.nlights, d.lights = len(order), lights
        d.cluster = 1 if spec.cluster == "power" else 0
        d.cam_pos[:], d.cam_dir[:], d.cam_up[:] = spec.cam_pos, spec.cam_dir, spec.cam_up
        d.fov_deg, d.sensor_width_mm = float(np.float32(spec.fov)), 36.0
        d.width, d.height, d.tile_x, d.tile_y = spec.width, spec.height, spec.tile, spec.tile
        d.filter = _FILTERS[spec.filter]
        d.filter_w = spec.filter_width if spec.filter_width is not None else _FILTER_W[spec.filter]
        d.filter_h = spec.filter_height if spec.filter_height is not None else _FILTER_W[spec.filter]
        fp = spec.filter_params
        third = float(np.float32(1.0) / np.float32(3.0))
        d.sigma_x, d.sigma_y = fp.get("sigma-x", 0.5), fp.get("sigma-y", 0.5)
        d.amplitude = fp.get("amplitude", 1.0)
        d.mitchell_b, d.mitchell_c = fp.get("b", third), fp.get("c", third)
        d.sampler = 1 if spec.sampler == "stratified" else 0
        d.spp, d.max_depth = spec.spp, spec.max_depth
        d.split_sah = 0 if spec.split_mode == "equal" else 1
        d.estimator = 1 if spec.estimator == "vpt" else 0
        out_name = spec.output.encode()
        d.output_filename = out_name
        h = C.c_void_p()
        ctx.check(ctx.L.mpt_scene_create(ctx.h, C.byref(d), C.byref(opts) if opts else None, C.byref(h)))
        del keep
        sc = cls(ctx, h)
        sc.tex_index = tex_index
        return sc

    def close(self):
        if getattr(self, "h", None):
            self.L.mpt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- inspection -----------------------------------------------------------------
    def stats(self) -> dict:
        st = SceneStats()
        self.ctx.check(self.L.mpt_scene_stats_get(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def binary_bvh(self) -> dict:
        n = C.c_int64()
        self.ctx.check(self.L.mpt_scene_binary_bvh(self.h, C.byref(n), None, None, None, None, None))
        nn = n.value
        b = np.zeros((nn, 6), np.float32)
        leaf = np.zeros(nn, np.int32)
        a = np.zeros(nn, np.int64)
        c = np.zeros(nn, np.int64)
        o = np.zeros(self.stats()["nprims"], np.int32)
        self.ctx.check(self.L.mpt_scene_binary_bvh(self.h, C.byref(n), b.ctypes.data, leaf.ctypes.data,
                                                   a.ctypes.data, c.ctypes.data, o.ctypes.data))
        return dict(bounds=b, leaf=leaf, a=a, b=c, ordered=o)

    def camera_matrices(self):
        f2c = np.zeros(16, np.float32)
        c2w = np.zeros(16, np.float32)
        self.ctx.check(self.L.mpt_scene_camera(self.h, f2c.ctypes.data, c2w.ctypes.data))
        return f2c.reshape(4, 4), c2w.reshape(4, 4)

    def light_radiance(self):
        n = self.stats()["nlights"]
        out = np.zeros((max(1, n), 3), np.float32)
        self.ctx.check(self.L.mpt_scene_light_radiance(self.h, out.ctypes.data))
        return out[:n]

    def light_pick_pdf(self):
        """Probability of picking each light (the scene's light cluster)."""
        n = self.stats()["nlights"]
        out = np.zeros(max(1, n), np.float32)
        self.ctx.check(self.L.mpt_scene_light_pick_pdf(self.h, out.ctypes.data))
        return out[:n]

    def last_stats(self) -> dict:
        st = Stats()
        self.ctx.check(self.L.mpt_stats_get(self.h, C.byref(st)))
        return st.as_dict()

    # -- Accelerator ------------------------------------------------------------------
    def trace(self, rays: np.ndarray, any_hit: bool = False):
        """rays: n x 8 float32 (ox oy oz tmin dx dy dz tmax), unit directions."""
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros((len(rays), 2), np.int32)
        self.ctx.check(self.L.mpt_trace(self.h, rays.ctypes.data, len(rays), int(any_hit), hits.ctypes.data))
        return hits[:, 0].copy(), hits[:, 1].view(np.float32).copy()

    def is_intersecting(self, rays):
        return self.trace(rays, False)

    def intersect_detail(self, rays: np.ndarray):
        """Closest hits with their surface detail: n x 14 float32 (primitive
        creation index, t, P, Ng, Ns, uvw); -1 and zeros for a miss."""
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((len(rays), 14), np.float32)
        self.ctx.check(self.L.mpt_intersect_detail(self.h, rays.ctypes.data, len(rays), out.ctypes.data))
        return out

    def is_occluded(self, rays):
        return self.trace(rays, True)[0].astype(bool)

    def trace_device(self, rays_ptr: int, n: int, hits_ptr: int, any_hit: bool = False, stream: int = 0):
        self.ctx.check(self.L.mpt_trace_device(self.h, rays_ptr, n, int(any_hit), hits_ptr, stream or None))

    # -- Renderer -----------------------------------------------------------------------
    def render(self, seed: int = 0, spp_begin: int = 0, spp_end: int = 0, max_depth: int = 0,
               accumulate: bool = False, readback: bool = True, keep_samples: bool = False, passwise: bool = False):
        p = RenderParams(seed, spp_begin, spp_end, max_depth, int(accumulate),
                         (MPT_RENDER_KEEP_SAMPLES if keep_samples else 0) | (MPT_RENDER_PASSWISE if passwise else 0))
        film = np.zeros((self.height, self.width, 4), np.float32) if readback else None
        self.ctx.check(self.L.mpt_render(self.h, C.byref(p), film.ctypes.data if readback else None))
        return film

    # -- differentiable rendering ---------------------------------------------------------
    def texture_info(self, index: int):
        """(width, height, channels, offset) of texture `index` in the flat texture array."""
        w, h, c, o = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self.ctx.check(self.L.mpt_scene_texture_info(self.h, index, C.byref(w), C.byref(h), C.byref(c), C.byref(o)))
        return w.value, h.value, c.value, o.value

    def get_texture(self, index: int) -> np.ndarray:
        """Texture `index` as the renderer samples it (H x W x C float32)."""
        w, h, c, _ = self.texture_info(index)
        out = np.zeros((h, w, c), np.float32)
        self.ctx.check(self.L.mpt_scene_get_texture(self.h, index, out.ctypes.data))
        return out

    def set_texture(self, index: int, data: np.ndarray):
        data = np.ascontiguousarray(data, np.float32)
        w, h, c, _ = self.texture_info(index)
        assert data.size == w * h * c, (data.shape, (h, w, c))
        self.ctx.check(self.L.mpt_scene_set_texture(self.h, index, data.ctypes.data))

    def set_light_radiance(self, light: int, rgb):
        self.ctx.check(self.L.mpt_scene_set_light_radiance(self.h, light,
                                                           np.ascontiguousarray(rgb, np.float32).ctypes.data))

    def adjoint(self, d_image: np.ndarray, seed: int = 0, spp_begin: int = 0, spp_end: int = 0, max_depth: int = 0):
        """dLoss/d(texture elements) (flat, float32) and dLoss/dLe (nlights x 3) for
        d_image = dLoss/dI (H x W x 3) of the render with the same arguments."""
        st = self.stats()
        d_image = np.ascontiguousarray(d_image, np.float32)
        assert d_image.shape == (self.height, self.width, 3)
        d_tex = np.zeros(max(1, st["tex_elements"]), np.float32)
        d_light = np.zeros((max(1, st["nlights"]), 3), np.float32)
        p = RenderParams(seed, spp_begin, spp_end, max_depth, 0, 0)
        self.ctx.check(self.L.mpt_render_adjoint(self.h, C.byref(p), d_image.ctypes.data, d_tex.ctypes.data,
                                                 d_light.ctypes.data))
        return d_tex[: st["tex_elements"]], d_light[: st["nlights"]]

    def texture_views(self, flat: np.ndarray):
        """Split a flat texture-element array into H x W x C views, one per texture."""
        out = []
        for i in range(self.stats()["ntextures"]):
            w, h, c, o = self.texture_info(i)
            out.append(flat[o:o + w * h * c].reshape(h, w, c))
        return out

    def film_device_ptr(self) -> int:
        p = C.c_void_p()
        self.ctx.check(self.L.mpt_film_device(self.h, C.byref(p)))
        return p.value

    def srgb8(self):
        out = np.zeros((self.height, self.width, 3), np.uint8)
        self.ctx.check(self.L.mpt_film_srgb8(self.h, out.ctypes.data))
        return out

    def save(self, path: str | None = None):
        """Film::save by extension; None: the film block's output-filename."""
        self.ctx.check(self.L.mpt_film_save(self.h, path.encode() if path is not None else None))

    def output_filename(self) -> str:
        n = C.c_size_t(0)
        self.ctx.check(self.L.mpt_scene_output_filename(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self.ctx.check(self.L.mpt_scene_output_filename(self.h, buf, n.value, None))
        return buf.value.decode()


def batch_opts(width: int = 4, stack_lds: int = 0, block: int = 128, capacity_log2: int = 21,
               graph_steps: int = 16, blocks_per_cu: int = 0) -> BatchOpts:
    """Zero stack_lds / blocks_per_cu: the library's default for the width."""
    return BatchOpts(width, stack_lds, block, capacity_log2, graph_steps, blocks_per_cu)


class Batch:
    """Many scenes rendered by one streaming wavefront (mpt_batch_*). The
    scenes must outlive the batch (it keeps references)."""

    def __init__(self, ctx: Context, scenes, opts: BatchOpts | None = None):
        self.ctx, self.L = ctx, ctx.L
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[s.h for s in self.scenes])
        h = C.c_void_p()
        ctx.check(self.L.mpt_batch_create(ctx.h, arr, len(self.scenes), C.byref(opts) if opts else None, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.mpt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _seeds(self, seeds):
        if seeds is None:
            return None
        a = np.ascontiguousarray(seeds, np.uint64)
        assert len(a) == len(self.scenes)
        return a

    def render(self, seeds=None, spp_override: int = 0) -> dict:
        sd = self._seeds(seeds)
        st = Stats()
        self.ctx.check(self.L.mpt_batch_render(self.h, sd.ctypes.data if sd is not None else None, spp_override,
                                               C.byref(st)))
        return st.as_dict()

    def film(self, i: int) -> np.ndarray:
        sc = self.scenes[i]
        out = np.zeros((sc.height, sc.width, 4), np.float32)
        self.ctx.check(self.L.mpt_batch_film(self.h, i, out.ctypes.data))
        return out

    @property
    def opts(self) -> dict:
        o = BatchOpts()
        self.ctx.check(self.L.mpt_batch_get_opts(self.h, C.byref(o)))
        return o.as_dict()

    def set_opts(self, **kw):
        o = BatchOpts(**{**self.opts, **kw})
        self.ctx.check(self.L.mpt_batch_set_opts(self.h, C.byref(o)))

    def autotune(self, seeds=None, trial_spp: int = 1, max_trials: int = 64):
        sd = self._seeds(seeds)
        best = BatchOpts()
        trials = (BatchTrial * max_trials)()
        nt = C.c_int32()
        self.ctx.check(self.L.mpt_batch_autotune(self.h, sd.ctypes.data if sd is not None else None, trial_spp,
                                                 C.byref(best), trials, max_trials, C.byref(nt)))
        return best.as_dict(), [(trials[i].opts.as_dict(), trials[i].ms) for i in range(min(nt.value, max_trials))]


def srgb8_device(ctx: Context, film_ptr: int, npix: int, out_ptr: int, stream: int = 0):
    """mpt_film_srgb8_device on device buffers (npix float4 in, npix x 3 bytes out)."""
    ctx.check(ctx.L.mpt_film_srgb8_device(ctx.h, film_ptr, npix, out_ptr, stream or None))


def image_write(path: str, rgb: np.ndarray):
    """mpt_image_write: 8-bit RGB (H x W x 3) to .png / .jpg / .ppm (host only)."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    assert rgb.ndim == 3 and rgb.shape[2] == 3
    L = native.load_library()
    rc = L.mpt_image_write(path.encode(), rgb.ctypes.data, rgb.shape[1], rgb.shape[0])
    if rc != 0:
        raise MptError(rc, f"mpt_image_write({path}) failed")


def film_mean(film: np.ndarray) -> np.ndarray:
    """Linear image sum(w * rgb) / sum(w) (Film::save without the transfer function)."""
    w = film[..., 3:4]
    return np.where(w > 0, film[..., :3] / np.where(w > 0, w, 1), 0).astype(np.float32)
