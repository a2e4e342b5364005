"""Pins the CPU oracle against checks that do not share its code: float64
library routines, an independent brute-force compositor (tests/_bruteforce.py),
closed forms, Python-integer key sorting, FP64 central finite differences,
textbook SH and torch.optim.Adam.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from sv_raster_amd import synthetic as S
from tests import _bruteforce as BF


def test_expf_ulp_bound():
    x = np.linspace(-103.0, 88.5, 1_000_001).astype(np.float32)
    y = O.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    m = ref > np.finfo(np.float32).tiny
    ulp = np.abs(y[m] - ref[m]) / np.spacing(ref[m].astype(np.float32)).astype(np.float64)
    assert ulp.max() <= 2.0
    special = O.expf(np.array([0.0, -np.inf, np.inf, 89.0, -104.0, np.nan], np.float32))
    assert special[0] == 1.0 and special[1] == 0.0 and np.isinf(special[2]) and np.isinf(special[3])
    assert special[4] == 0.0 and np.isnan(special[5])


def _small_case(seed, name="C", direction=None, W=64, H=64):
    sc = S.make_scene(name, seed=seed)
    cam = S.make_camera(W, H, seed=seed, direction=direction)
    return sc, cam


@pytest.mark.parametrize("seed,direction", [(0, None), (1, None), (3, (0.02, 0.03, 1.0))])
def test_preprocess_bbox_contains_every_hit_pixel(seed, direction):
    sc, cam = _small_case(seed, direction=direction)
    n_dup, bbox, qbits = O.preprocess(cam, sc["vox_centers"], sc["vox_lengths"])
    W, H = cam["W"], cam["H"]
    rng = np.random.default_rng(seed)
    c2w = cam["c2w"].astype(np.float64)
    w2c = cam["w2c"].astype(np.float64)
    ro = c2w[:3, 3]
    checked = 0
    for v in rng.choice(len(n_dup), 150, replace=False):
        c = sc["vox_centers"][v].astype(np.float64)
        l = float(sc["vox_lengths"][v, 0])
        corners = c + 0.5 * l * np.array([[(i >> 2 & 1) * 2 - 1, (i >> 1 & 1) * 2 - 1, (i & 1) * 2 - 1]
                                          for i in range(8)])
        cam_z = (w2c[:3, :3] @ corners.T + w2c[:3, 3:4])[2]
        if (cam_z < cam["near"] + 1e-3).any():
            continue
        # quadrant bits: octant of each corner relative to the camera centre
        q = ((corners[:, 0] < ro[0]) << 2) | ((corners[:, 1] < ro[1]) << 1) | (corners[:, 2] < ro[2])
        assert int(qbits[v]) == int(np.bitwise_or.reduce(1 << q.astype(np.int64)))
        bx, by = int(bbox[v, 0]), int(bbox[v, 1])
        x0, y0, x1, y1 = bx >> 16, bx & 0xFFFF, by >> 16, by & 0xFFFF
        for py in range(H):
            for px in range(W):
                o, d = BF.ray(px, py, cam)
                a, b = BF.slab(c, l, o, d)
                if a > b + 1e-6 * max(1.0, abs(b)) or abs(a - b) < 1e-6:
                    continue
                if a > b:
                    continue
                assert x0 <= px <= x1 and y0 <= py <= y1, (v, px, py, (x0, y0, x1, y1))
        if n_dup[v]:
            gx, gy = (W + 15) // 16, (H + 15) // 16
            tiles = (min(y1 // 16, gy - 1) - min(y0 // 16, gy - 1) + 1) * (min(x1 // 16, gx - 1) - min(x0 // 16, gx - 1) + 1)
            assert n_dup[v] == tiles * bin(int(qbits[v])).count("1")
        checked += 1
    assert checked > 50


@pytest.mark.parametrize("seed,direction", [(0, None), (3, (0.02, 0.03, 1.0)), (4, (1.0, -0.01, 0.02))])
def test_binning_matches_python_int_sort(seed, direction):
    sc, cam = _small_case(seed, direction=direction)
    n_dup, bbox, qbits = O.preprocess(cam, sc["vox_centers"], sc["vox_lengths"])
    hi, lo, vl, ranges = O.binning(cam, sc["octpath"], n_dup, bbox, qbits)
    lists = BF.tile_lists(cam, sc["octpath"].reshape(-1), n_dup, bbox, qbits)
    # python-int 128-bit keys, stable sort = sorted() with the emission index as tie-break
    ent = []
    for tile, items in lists.items():
        for rank, v, q in items:
            ent.append(((tile << 63) | rank, v, q))
    ent.sort()
    keys = np.array([(k >> 64, k & ((1 << 64) - 1)) for k, _, _ in ent], dtype=np.uint64)
    assert np.array_equal(hi, keys[:, 0]) and np.array_equal(lo, keys[:, 1])
    assert np.array_equal(vl, np.array([(q << 29) | v for _, v, q in ent], np.uint32))
    tiles = (keys[:, 0].astype(object) * 2 + (keys[:, 1] >> np.uint64(63)).astype(object)).astype(np.int64)
    T = ranges.shape[0]
    for t in range(T):
        s, e = np.searchsorted(tiles, t, "left"), np.searchsorted(tiles, t, "right")
        exp = (s, e) if e > s else (0, 0)
        assert tuple(ranges[t]) == exp


@pytest.mark.parametrize("seed,direction", [(0, None), (2, None), (3, (0.02, 0.03, 1.0))])
def test_forward_matches_bruteforce(seed, direction):
    sc, cam = _small_case(seed, direction=direction)
    rgbs = O.sh_fwd(3, None, sc["vox_centers"], cam["c2w"][:3, 3], sc["sh0"], sc["shs"])
    out = O.full_forward(cam, sc["octpath"], sc["vox_centers"], sc["vox_lengths"], sc["geos"], rgbs)
    color, T, nc, amb = BF.render(cam, sc["octpath"].reshape(-1), out["n_dup"], out["bbox"], out["qbits"],
                                  sc["vox_centers"], sc["vox_lengths"].reshape(-1), sc["geos"], rgbs)
    ok = ~amb
    assert ok.mean() > 0.95
    assert np.abs(out["color"][:, ok] - color[:, ok]).max() < 2e-5
    assert np.abs(out["T"][ok] - T[ok]).max() < 2e-5
    assert np.array_equal(out["n_contrib"][ok], nc[ok])
    gx = (cam["W"] + 15) // 16
    for t in range(out["ranges"].shape[0]):
        ty, tx = divmod(t, gx)
        tile_nc = out["n_contrib"][ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16]
        assert out["tile_last"][t] == out["ranges"][t, 0] + tile_nc.max()


def test_single_voxel_closed_form():
    # one voxel of constant corner density seen straight through its centre
    cam = S.make_camera(32, 32, seed=0, direction=(0.0, 0.0, -1.0))
    for d, l in ((-1.0, 0.25), (0.5, 0.125), (2.0, 0.03125)):
        centers = np.zeros((1, 3), np.float32)
        lengths = np.full((1, 1), l, np.float32)
        geos = np.full((1, 8), d, np.float32)
        rgbs = np.array([[0.25, 0.5, 0.75]], np.float32)
        out = O.full_forward(cam, np.zeros((1, 1), np.int64), centers, lengths, geos, rgbs)
        el = d if d > 1.1 else math.exp(0.909090909091 * d - 0.904689820196)
        for px, py in ((15, 15), (16, 16), (15, 16)):
            o, ddir = BF.ray(px, py, cam)
            a, b = BF.slab(np.zeros(3), l, o, ddir)
            if a > b:  # this pixel's ray misses the smallest voxel
                assert out["T"][py, px] == 1.0 and not out["color"][:, py, px].any()
                continue
            alpha = min(0.99999, 1.0 - math.exp(-100.0 * (b - a) * el))
            assert abs(out["T"][py, px] - (1 - alpha)) < 1e-6
            assert np.allclose(out["color"][:, py, px], alpha * rgbs[0], atol=1e-6)
            # the voxel straddles the camera's x / y planes, so the tile list holds one
            # copy per quadrant; n_contrib is the list position of this pixel's copy
            k = int(out["n_contrib"][py, px]) - 1
            assert k >= 0 and (int(out["vox_list"][k]) >> 29) == BF.quad_of(ddir)


def _fd_scene(seed=0):
    sc = S.make_scene("C", seed=seed)
    rng = np.random.default_rng(seed + 5)
    # moderate densities keep every alpha away from the MAX_ALPHA clamp
    sc["geos"] = rng.normal(-2.0, 0.5, sc["geos"].shape)
    cam = S.make_camera(64, 64, seed=seed)
    return sc, cam


def _loss64(cam, sc, base, geos, rgbs, wc, wt):
    color, T, nc, _ = O.render_fwd(cam, base["ranges"], base["vox_list"], base["bbox"], sc["vox_centers"],
                                   sc["vox_lengths"], geos, rgbs, dtype=np.float64)
    return float((wc * color).sum() + (wt * T).sum()), nc


def _rgb64(sc, cam, sh0, shs, idx):
    rgbs = np.zeros((sc["vox_centers"].shape[0], 3))
    bas = {}
    for v in idx:
        rgbs[v], bas[v] = BF.sh_rgb(sc["vox_centers"][v], cam["c2w"][:3, 3], sh0[v], shs[v])
    return rgbs, bas


def test_backward_fp64_central_differences():
    """200 random density and SH parameters, eps 1e-6, vs the oracle's FP64 adjoint."""
    sc, cam = _fd_scene()
    n_dup, bbox, qbits = O.preprocess(cam, sc["vox_centers"], sc["vox_lengths"])
    _, _, vl, ranges = O.binning(cam, sc["octpath"], n_dup, bbox, qbits)
    base = dict(ranges=ranges, vox_list=vl, bbox=bbox)
    idx = np.nonzero(n_dup)[0]
    sh0 = sc["sh0"].astype(np.float64)
    shs = sc["shs"].astype(np.float64)
    geos = sc["geos"].astype(np.float64)
    rgbs, basis = _rgb64(sc, cam, sh0, shs, idx)
    assert (rgbs[idx] > 0).all()  # SH clamp inactive (the reference backward ignores it)
    rng = np.random.default_rng(11)
    H, W = cam["H"], cam["W"]
    wc = rng.normal(size=(3, H, W))
    wt = rng.normal(size=(H, W))
    L0, nc0 = _loss64(cam, sc, base, geos, rgbs, wc, wt)
    color, T, nc, tl = O.render_fwd(cam, ranges, vl, bbox, sc["vox_centers"], sc["vox_lengths"], geos, rgbs,
                                    dtype=np.float64)
    dg, dr, _ = O.render_bwd(cam, ranges, vl, bbox, sc["vox_centers"], sc["vox_lengths"], geos, rgbs, T, nc, tl,
                             wc, wt, dtype=np.float64)
    touched = np.unique(vl & 0x1FFFFFFF)
    eps = 1e-6
    checked, bad = 0, []
    for k in range(200):
        v = int(rng.choice(touched))
        if k % 2 == 0:
            i = int(rng.integers(8))
            p = geos.copy(); p[v, i] += eps
            Lp, ncp = _loss64(cam, sc, base, p, rgbs, wc, wt)
            p[v, i] -= 2 * eps
            Lm, ncm = _loss64(cam, sc, base, p, rgbs, wc, wt)
            ana = dg[v, i]
        else:
            c, ch = int(rng.integers(16)), int(rng.integers(3))
            s0, s = sh0.copy(), shs.copy()
            (s0 if c == 0 else s)[(v, ch) if c == 0 else (v, c - 1, ch)] += eps
            rp, _ = _rgb64(sc, cam, s0, s, [v]); r2 = rgbs.copy(); r2[v] = rp[v]
            Lp, ncp = _loss64(cam, sc, base, geos, r2, wc, wt)
            (s0 if c == 0 else s)[(v, ch) if c == 0 else (v, c - 1, ch)] -= 2 * eps
            rm, _ = _rgb64(sc, cam, s0, s, [v]); r2 = rgbs.copy(); r2[v] = rm[v]
            Lm, ncm = _loss64(cam, sc, base, geos, r2, wc, wt)
            ana = basis[v][c] * dr[v, ch]
        if not (np.array_equal(ncp, nc0) and np.array_equal(ncm, nc0)):
            continue  # a threshold decision flipped: not differentiable there
        fd = (Lp - Lm) / (2 * eps)
        checked += 1
        if abs(fd - ana) > 1e-4 * max(abs(fd), abs(ana)) + 1e-7:
            bad.append((k, v, fd, ana))
    assert checked >= 180, checked
    assert not bad, bad[:5]


def test_sh_against_textbook_basis():
    sc = S.make_scene("C", seed=0)
    cam_pos = np.array([0.3, -2.0, 1.5], np.float32)
    rgbs = O.sh_fwd(3, None, sc["vox_centers"], cam_pos, sc["sh0"], sc["shs"])
    g = np.random.default_rng(0).normal(size=rgbs.shape).astype(np.float32)
    d0, ds = O.sh_bwd(3, 16, None, sc["vox_centers"], cam_pos, g)
    for v in range(0, sc["vox_centers"].shape[0], 37):
        ref, b = BF.sh_rgb(sc["vox_centers"][v], cam_pos, sc["sh0"][v], sc["shs"][v])
        assert np.allclose(rgbs[v], ref, atol=2e-6)
        assert np.allclose(d0[v], b[0] * g[v], rtol=1e-5, atol=1e-7)
        assert np.allclose(ds[v], b[1:, None] * g[v][None, :], rtol=1e-5, atol=1e-7)


def test_adam_matches_torch_adam():
    import torch
    rng = np.random.default_rng(0)
    p = rng.normal(size=4096).astype(np.float32)
    m = np.zeros_like(p); v = np.zeros_like(p)
    tp = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=0.025, betas=(0.1, 0.99), eps=1e-15)
    for step in range(1, 6):
        g = rng.normal(size=p.shape).astype(np.float32)
        O.adam_step(p, g, m, v, step, 0.025, 0.1, 0.99, 1e-15)
        tp.grad = torch.tensor(g)
        opt.step()
    assert np.allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-6)


def test_sparse_adam_skips_zero_gradients():
    rng = np.random.default_rng(1)
    p = rng.normal(size=1000).astype(np.float32)
    m = rng.normal(size=1000).astype(np.float32) * 0.1
    v = np.abs(rng.normal(size=1000)).astype(np.float32) * 0.1
    g = rng.normal(size=1000).astype(np.float32)
    g[::3] = 0.0
    p0, m0, v0 = p.copy(), m.copy(), v.copy()
    O.adam_step(p, g, m, v, 4, 0.01, 0.1, 0.99, 1e-15, sparse=True)
    z = g == 0
    assert np.array_equal(p[z], p0[z]) and np.array_equal(m[z], m0[z]) and np.array_equal(v[z], v0[z])
    assert (p[~z] != p0[~z]).all()
    # closed form of one unbiased step on the touched elements
    mm = 0.1 * m0 + 0.9 * g
    vv = 0.99 * v0 + 0.01 * g * g
    step = 0.01 / (1 - 0.1 ** 4) * mm / (np.sqrt(vv) / math.sqrt(1 - 0.99 ** 4) + 1e-15)
    assert np.allclose(p[~z], (p0 - step)[~z], rtol=1e-5, atol=1e-7)


def test_octree_codec_and_order_rank():
    tree = S.make_octree(3000, 3, 7, seed=2)
    lv = tree["octlevel"].reshape(-1)
    ijk = S.octpath_to_ijk(tree["octpath"], lv)
    assert np.array_equal(S.ijk_to_octpath(ijk, lv), tree["octpath"].reshape(-1))
    # hand interleave of the first voxel (x high bit of each 3-bit group, coarsest level first)
    i, j, k = (int(a) for a in ijk[5]); L = int(lv[5])
    path = 0
    for l in range(L - 1, -1, -1):
        path = (path << 3) | (((i >> l) & 1) << 2) | (((j >> l) & 1) << 1) | ((k >> l) & 1)
    assert path << (3 * (21 - L)) == int(tree["octpath"][5, 0])
    ranks = O.voxel_order_rank(tree["octpath"])
    for q in range(8):
        rep = int("0o" + str(q) * 21, 8)
        assert all(int(r) == (int(p) ^ rep) for r, p in zip(ranks[q][:50], tree["octpath"][:50, 0]))


@pytest.mark.parametrize("name", ["C", "A"])
def test_generator_is_a_partition(name):
    cfg = S.CONFIGS[name]
    tree = S.make_octree(cfg["n"], cfg["lv_min"], cfg["lv_max"], seed=0, outside_level=cfg["outside"])
    lv = tree["octlevel"].reshape(-1).astype(np.int64)
    assert lv.min() >= cfg["lv_min"] and lv.max() <= cfg["lv_max"]
    assert abs(len(lv) - cfg["n"]) <= 7
    assert np.isclose((8.0 ** -lv).sum(), 1.0)          # leaves tile the unit cube
    assert len(np.unique(tree["octpath"])) == len(lv)
    again = S.make_octree(cfg["n"], cfg["lv_min"], cfg["lv_max"], seed=0, outside_level=cfg["outside"])
    assert np.array_equal(again["octpath"], tree["octpath"])


def test_shell_layout_covers_outside_region():
    tree = S.make_octree(5000, 3, 6, seed=0, outside_level=5)
    lv = tree["octlevel"].reshape(-1).astype(np.int64)
    # scene extent 32: total volume of all leaves is 32^3 (shell + inner unit cube)
    assert np.isclose((32.0 ** 3 * 8.0 ** -lv).sum(), 32.0 ** 3)
    assert tree["is_shell"].sum() == 56 * 5
