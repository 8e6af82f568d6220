"""Shared test helpers: seeded ray sets and scene construction."""
import numpy as np

from cadise_amd import scenes
from oracle.binding import OracleScene, oracle_lib

FLT_MAX = float(np.finfo(np.float32).max)


def normalize_ref(d):
    """Directions normalized exactly as the reference's TVector::normalize."""
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros_like(d)
    oracle_lib().or_normalize(d, len(d), out)
    return out


def stable_dirs(d):
    """Directions that the reference's normalize maps to themselves bit for bit
    (its Ray constructor normalizes every direction it is given), by
    re-normalizing until a fixed point; rows that never settle are dropped."""
    d = normalize_ref(d)
    keep = np.ones(len(d), bool)
    for _ in range(4):
        nd = normalize_ref(d)
        same = np.all(nd.view(np.int32) == d.view(np.int32), axis=1)
        d = nd
        keep = same
    return d, keep


def random_rays(n, lo, hi, seed, tmin=1e-3):
    """n x 8 rays: origins uniform in the box [lo, hi], directions isotropic."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d, keep = stable_dirs(rng.normal(size=(n, 3)).astype(np.float32))
    d[~keep] = normalize_ref(d[~keep])
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = FLT_MAX
    return r


def camera_rays(spec, n, seed, oracle=None):
    """Primary rays of the scene camera at random film positions (reference-exact)."""
    rng = np.random.default_rng(seed)
    fxy = np.stack([rng.uniform(0, spec.width, n), rng.uniform(0, spec.height, n)], 1).astype(np.float32)
    od = (oracle or OracleScene(spec)).camera_rays(fxy)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = od[:, :3]
    r[:, 3] = 1e-3
    r[:, 4:7] = od[:, 3:]
    r[:, 7] = FLT_MAX
    return r


def small_spheres(n_spheres=8, level=4, **kw):
    kw.setdefault("width", 64)
    kw.setdefault("height", 48)
    kw.setdefault("spp", 4)
    kw.setdefault("offset", (0.0, 0.0, 0.0))
    return scenes.procedural_spheres(n_spheres=n_spheres, level=level, **kw)


def stress_rays(spec, oracle, n, seed):
    """n seeded rays for the traversal stress check: half camera primaries,
    half secondary rays leaving random surface points with cosine-distributed
    directions (directions at the reference's normalize fixed point)."""
    cam = camera_rays(spec, n // 2, seed, oracle)
    need = n - n // 2
    sec = scenes.surface_rays(spec, int(need * 1.25) + 1000, seed + 1)
    d, keep = stable_dirs(sec[:, 4:7])
    sec[:, 4:7] = d
    sec = sec[keep][:need]
    assert len(sec) == need
    return np.concatenate([cam, sec])
