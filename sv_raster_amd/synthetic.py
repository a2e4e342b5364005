"""Seeded synthetic sparse-voxel scenes and cameras (numpy only).

Scenes follow the reference's octree conventions:
  * octpaths are 21-level Morton paths, coarsest level in the top bits, x the
    high bit of each 3-bit group (src/sv_raster/new/utils/octree_utils.py:59-134,
    backends/new_cuda/src/utils.cu:129-189);
  * the optional outside shell is the reference's ``gen_octpath_shell(lv, 1)``
    for lv = 1..outside_level (octree_utils.py:262-291, constructor.py:464-475):
    56 voxels per shell level, scene extent 2**outside_level, and the inner unit
    cube [-0.5, 0.5]^3 at absolute level outside_level + 1;
  * voxel centres / sizes are decoded as in octree_utils.py:107-134.

Inner leaves come from a dense grid at ``lv_min`` that is refined, level by
level, on cells crossing the surface of one of 8 random spheres (centres
U(-0.5, 0.5)^3, radius U(0.05, 0.3)); at each level a random subset of those
cells is split so that the leaf budget is spread down to ``lv_max``.

Cameras sit on a radius-3 sphere looking at the origin with the COLMAP
convention (x right, y down, z forward) built like gen_circular_poses
(src/sv_raster/new/utils/camera_utils.py:61-79).
"""
import math

import numpy as np

MAX_NUM_LEVELS = 21
SHELL_GEO_MEAN = -10.0   # the reference's geo_init (src/sv_raster/new/config.py:188)


def ijk_to_octpath(ijk, level):
    """Morton-encode integer voxel coordinates at the given levels (utils.cu:129-158)."""
    ijk = np.asarray(ijk, dtype=np.int64).reshape(-1, 3)
    level = np.asarray(level, dtype=np.int64).reshape(-1)
    path = np.zeros(ijk.shape[0], dtype=np.uint64)
    i, j, k = (ijk[:, 0].astype(np.uint64), ijk[:, 1].astype(np.uint64), ijk[:, 2].astype(np.uint64))
    one = np.uint64(1)
    for lv in range(MAX_NUM_LEVELS):
        s = np.uint64(lv)
        bits = (((i >> s) & one) << np.uint64(2)) | (((j >> s) & one) << one) | ((k >> s) & one)
        path |= bits << np.uint64(3 * lv)
    path <<= (3 * (MAX_NUM_LEVELS - level)).astype(np.uint64)
    return path.astype(np.int64)


def octpath_to_ijk(octpath, level):
    """Inverse of ijk_to_octpath (utils.cu:160-189)."""
    path = np.asarray(octpath, dtype=np.int64).reshape(-1).astype(np.uint64)
    level = np.asarray(level, dtype=np.int64).reshape(-1)
    path = path >> (3 * (MAX_NUM_LEVELS - level)).astype(np.uint64)
    ijk = np.zeros((path.size, 3), dtype=np.int64)
    one = np.uint64(1)
    for lv in range(MAX_NUM_LEVELS):
        bits = (path >> np.uint64(3 * lv)) & np.uint64(7)
        ijk[:, 0] |= (((bits >> np.uint64(2)) & one).astype(np.int64)) << lv
        ijk[:, 1] |= (((bits >> one) & one).astype(np.int64)) << lv
        ijk[:, 2] |= ((bits & one).astype(np.int64)) << lv
    return ijk


def shell_octpaths(shell_level):
    """The reference's gen_octpath_shell(shell_level, n_level_inside=1): 56 voxels at level shell_level+1."""
    eight = np.arange(8, dtype=np.int64)
    path = eight << (3 * (MAX_NUM_LEVELS - 1))
    for k in range(shell_level - 1):
        path |= (eight ^ 7) << (3 * (MAX_NUM_LEVELS - (k + 2)))
    path = path.reshape(8, 1) | (eight << (3 * (MAX_NUM_LEVELS - shell_level - 1))).reshape(1, 8)
    keep = eight.reshape(1, 8) != (eight ^ 7).reshape(8, 1)
    path = path[keep]
    return path, np.full(path.shape, shell_level + 1, dtype=np.int64)


def _crosses_surface(centers, half, sph_c, sph_r):
    hit = np.zeros(centers.shape[0], dtype=bool)
    for s in range(sph_c.shape[0]):
        d = np.abs(centers - sph_c[s])
        dmin = np.sqrt((np.maximum(d - half, 0.0) ** 2).sum(1))
        dmax = np.sqrt(((d + half) ** 2).sum(1))
        hit |= (dmin <= sph_r[s]) & (sph_r[s] <= dmax)
    return hit


_CHILD = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], dtype=np.int64)


def make_octree(n_target, lv_min, lv_max, seed, outside_level=0):
    """Return dict(octpath int64 [N,1], octlevel int8 [N,1], is_shell bool [N], scene_center, scene_extent)."""
    rng = np.random.default_rng(seed)
    sph_c = rng.uniform(-0.5, 0.5, (8, 3))
    sph_r = rng.uniform(0.05, 0.3, 8)
    n_shell = 56 * outside_level
    budget = n_target - n_shell
    g = 2 ** lv_min
    ax = np.arange(g, dtype=np.int64)
    front = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    count = front.shape[0]
    leaves_ijk, leaves_lv = [], []
    L = lv_min
    while L < lv_max:
        h = 1.0 / (2 ** L)
        centers = -0.5 + (front + 0.5) * h
        cand = np.nonzero(_crosses_surface(centers, 0.5 * h, sph_c, sph_r))[0]
        kmax = (budget - count) // 7
        if kmax <= 0 or cand.size == 0:
            break
        share = kmax if L == lv_max - 1 else -(-kmax // (lv_max - L))
        k = min(cand.size, share)
        chosen = np.sort(rng.choice(cand, size=k, replace=False))
        split = np.zeros(front.shape[0], dtype=bool)
        split[chosen] = True
        leaves_ijk.append(front[~split])
        leaves_lv.append(np.full((~split).sum(), L, dtype=np.int64))
        front = (front[split] * 2)[:, None, :] + _CHILD[None, :, :]
        front = front.reshape(-1, 3)
        count += 7 * k
        L += 1
    leaves_ijk.append(front)
    leaves_lv.append(np.full(front.shape[0], L, dtype=np.int64))
    ijk = np.concatenate(leaves_ijk)
    lv = np.concatenate(leaves_lv)
    # inner-cube coordinates -> absolute octree coordinates
    if outside_level > 0:
        offset = (2 ** (outside_level - 1) * 2 - 1) * (2 ** lv) // 2   # (2^(out-1) - 0.5) * 2^L
        ijk = ijk + offset[:, None]
        lv = lv + outside_level
    paths = [ijk_to_octpath(ijk, lv)]
    levels = [lv]
    shell = [np.zeros(lv.size, dtype=bool)]
    for s in range(1, outside_level + 1):
        p, l = shell_octpaths(s)
        paths.append(p)
        levels.append(l)
        shell.append(np.ones(p.size, dtype=bool))
    octpath = np.concatenate(paths)
    octlevel = np.concatenate(levels)
    is_shell = np.concatenate(shell)
    order = np.argsort(octpath, kind="stable")
    return dict(
        octpath=octpath[order].reshape(-1, 1),
        octlevel=octlevel[order].astype(np.int8).reshape(-1, 1),
        is_shell=is_shell[order],
        scene_center=np.zeros(3, dtype=np.float32),
        scene_extent=np.float32(2.0 ** outside_level),
    )


def decode(octpath, octlevel, scene_center, scene_extent):
    """Voxel centres [N,3] and sizes [N,1] in float32 (octree_utils.py:59-68,107-134)."""
    lv = np.asarray(octlevel, dtype=np.int64).reshape(-1)
    ijk = octpath_to_ijk(octpath, lv)
    size = np.ldexp(np.float32(scene_extent), -lv).astype(np.float32).reshape(-1, 1)
    scene_min = (np.asarray(scene_center, np.float32) - np.float32(0.5) * np.float32(scene_extent)).astype(np.float32)
    centers = (ijk.astype(np.float32) + np.float32(0.5)) * size + scene_min
    return centers.astype(np.float32), size


def make_params(n, seed, is_shell=None, sh_degree=3):
    """Per-voxel trainable parameters: geos [N,8] N(-1,1) (shell: N(-10,1)), sh0 U(0,1), shs N(0,0.05)."""
    rng = np.random.default_rng(seed + 7919)
    geos = rng.normal(-1.0, 1.0, (n, 8)).astype(np.float32)
    if is_shell is not None and is_shell.any():
        geos[is_shell] = rng.normal(SHELL_GEO_MEAN, 1.0, (int(is_shell.sum()), 8)).astype(np.float32)
    sh0 = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    shs = rng.normal(0.0, 0.05, (n, (sh_degree + 1) ** 2 - 1, 3)).astype(np.float32)
    return dict(geos=geos, sh0=sh0, shs=shs)


def make_camera(W, H, seed, radius=3.0, fovx_deg=50.0, near=0.02, direction=None):
    """Pinhole camera on a sphere looking at the origin; matrices float32 row-major 4x4.
    `direction` (unit vector from the origin to the camera) overrides the seeded draw."""
    rng = np.random.default_rng(seed + 104729)
    while direction is None:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if abs(d[1]) < 0.95:
            direction = d
    d = np.asarray(direction, np.float64)
    pos = radius * d / np.linalg.norm(d)
    lookat = -pos / np.linalg.norm(pos)
    down = np.array([0.0, 1.0, 0.0])
    right = np.cross(down, lookat)
    right /= np.linalg.norm(right)
    down = np.cross(lookat, right)
    c2w = np.eye(4)
    c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = right, down, lookat, pos
    w2c = np.linalg.inv(c2w)
    tanfovx = math.tan(math.radians(fovx_deg) * 0.5)
    tanfovy = tanfovx * H / W
    return dict(W=int(W), H=int(H), tanfovx=float(np.float32(tanfovx)), tanfovy=float(np.float32(tanfovy)),
                cx=float(W) / 2.0, cy=float(H) / 2.0, near=float(near),
                c2w=c2w.astype(np.float32), w2c=w2c.astype(np.float32))


# Workloads (octree size, levels, shell, views).
CONFIGS = {
    "A": dict(n=100_000, lv_min=5, lv_max=8, outside=0, views=[(800, 800)]),
    "B": dict(n=1_000_000, lv_min=4, lv_max=12, outside=5, views=[(1920, 1080)]),
    "C": dict(n=2_000, lv_min=3, lv_max=6, outside=0, views=[(64, 64)]),
    "D": dict(n=2_000_000, lv_min=4, lv_max=14, outside=0, views="mixed16"),
    "E": dict(n=8_000_000, lv_min=4, lv_max=16, outside=5, views="mixed64"),
}

_RES_D = [(640, 480), (1280, 720), (1920, 1080), (2560, 1440)]
_RES_E = [(1280, 720), (1920, 1080), (2560, 1440)]


def config_views(name, seed=0):
    cfg = CONFIGS[name]
    if isinstance(cfg["views"], list):
        return cfg["views"]
    rng = np.random.default_rng(seed + 31337)
    if cfg["views"] == "mixed16":
        return [_RES_D[i] for i in rng.integers(0, len(_RES_D), 16)]
    return [_RES_E[i] for i in rng.integers(0, len(_RES_E), 64)]


def make_scene(name_or_cfg, seed=0, n=None):
    """Octree + decoded geometry + parameters for a named workload (or a dict with the same keys)."""
    cfg = CONFIGS[name_or_cfg] if isinstance(name_or_cfg, str) else name_or_cfg
    n = cfg["n"] if n is None else n
    tree = make_octree(n, cfg["lv_min"], cfg["lv_max"], seed, cfg["outside"])
    centers, sizes = decode(tree["octpath"], tree["octlevel"], tree["scene_center"], tree["scene_extent"])
    params = make_params(centers.shape[0], seed, tree["is_shell"])
    return dict(tree, vox_centers=centers, vox_lengths=sizes, **params)


def make_target(H, W, seed):
    return np.random.default_rng(seed + 65537).uniform(0.0, 1.0, (3, H, W)).astype(np.float32)
