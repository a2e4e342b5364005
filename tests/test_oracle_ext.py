"""The oracle's compositor options (depth, distortion, normal, max_w) and the
regularizer gradients (R_concen, ascending, distortion) pinned without the
reference: the forward against an independent float64 brute force, the
backward of every differentiable output by FP64 central differences, and the
regularizers whose reference gradient is partial (R_concen: colours only;
ascending: densities at fixed weights) by their closed form over the brute
force's per-pixel entries."""
import numpy as np
import pytest

from oracle import oracle as O
from sv_raster_amd import synthetic as S
from tests import _bruteforce as BF
from tests._harness import make_case, run_oracle


def _small_case(seed, W=40, H=32):
    cfg = dict(S.CONFIGS["C"], views=[(W, H)])
    case = make_case("C", seed=seed, cfg=cfg)
    o = run_oracle(case, backward=False)
    return case, o


def _ext(case, o, dtype=np.float32, n_samp=1, geos=None, rgbs=None, **kw):
    sc, cam = case["scene"], case["cam"]
    return O.render_fwd_ext(cam, o["ranges"], o["vox_list"], o["bbox"], sc["vox_centers"], sc["vox_lengths"],
                            sc["geos"] if geos is None else geos, o["rgbs"] if rgbs is None else rgbs,
                            n_samp=n_samp, dtype=dtype, **kw)


def test_ext_with_options_off_is_the_plain_compositor():
    case, o = _small_case(0)
    e = _ext(case, o)
    assert np.array_equal(e["color"], o["color"]) and np.array_equal(e["T"], o["T"])
    assert np.array_equal(e["n_contrib"], o["n_contrib"]) and np.array_equal(e["tile_last"], o["tile_last"])
    e2 = _ext(case, o, need_depth=True, need_dist=True, need_normal=True, track_max_w=True)
    assert np.array_equal(e2["color"], o["color"]) and np.array_equal(e2["n_contrib"], o["n_contrib"])


@pytest.mark.parametrize("seed", [1, 2])
def test_ext_forward_matches_bruteforce(seed):
    case, o = _small_case(seed)
    sc, cam = case["scene"], case["cam"]
    W, H = cam["W"], cam["H"]
    e = _ext(case, o, need_depth=True, need_dist=True, need_normal=True, track_max_w=True)
    pix = [(x, y) for y in range(H) for x in range(W)]
    bf, mw = BF.render_ext(cam, sc["octpath"], o["n_dup"], o["bbox"], o["qbits"], sc["vox_centers"],
                           sc["vox_lengths"], sc["geos"], o["rgbs"], pix)
    _, _, _, amb = BF.render(cam, sc["octpath"], o["n_dup"], o["bbox"], o["qbits"], sc["vox_centers"],
                             sc["vox_lengths"], sc["geos"], o["rgbs"])
    checked = 0
    for (x, y), r in bf.items():
        if amb[y, x]:
            continue
        checked += 1
        for got, ref in ((e["depth"][0, y, x], r["depth"]), (e["depth"][2, y, x], r["median"]),
                         (e["depth"][1, y, x], r["dist"])):
            assert abs(got - ref) <= 1e-4 * max(1.0, abs(ref)), (x, y, got, ref)
        assert np.abs(e["normal"][:, y, x] - r["normal"]).max() <= 1e-4 * max(1.0, np.abs(r["normal"]).max())
    assert checked > 0.9 * len(pix)
    assert np.abs(e["max_w"] - mw).max() <= 1e-4  # float32 vs float64 weights


def _fd_setup(seed, n_samp):
    sc = S.make_scene("C", seed=seed)
    rng = np.random.default_rng(seed + 5)
    sc["geos"] = rng.normal(-2.0, 0.5, sc["geos"].shape)
    cam = S.make_camera(48, 40, seed=seed)
    n_dup, bbox, qbits = O.preprocess(cam, sc["vox_centers"], sc["vox_lengths"])
    _, _, vl, ranges = O.binning(cam, sc["octpath"], n_dup, bbox, qbits)
    o = dict(ranges=ranges, vox_list=vl, bbox=bbox)
    rgbs = rng.uniform(0.1, 0.9, (sc["vox_centers"].shape[0], 3))
    return sc, cam, o, rgbs, rng


@pytest.mark.parametrize("n_samp", [1, 3])
def test_ext_backward_fp64_central_differences(n_samp):
    """L = <wc, color> + <wt, T> + <wd, mean depth> + <wn, normal> + lambda_dist / (W H) * sum of the
    per-pixel distortion loss; 160 density and colour parameters, eps 1e-6."""
    sc, cam, o, rgbs, rng = _fd_setup(3, n_samp)
    W, H = cam["W"], cam["H"]
    geos = sc["geos"].astype(np.float64)
    case = dict(scene=sc, cam=cam)
    wc, wt = rng.normal(size=(3, H, W)), rng.normal(size=(H, W))
    wd = np.zeros((3, H, W)); wd[0] = rng.normal(size=(H, W))
    wn = rng.normal(size=(3, H, W))
    lam = 0.7

    def loss(g, r):
        e = _ext(case, o, np.float64, n_samp, g, r, need_depth=True, need_dist=True, need_normal=True,
                 dist_loss=True)
        L = (wc * e["color"]).sum() + (wt * e["T"]).sum() + (wd * e["depth"]).sum() + (wn * e["normal"]).sum()
        return float(L + lam / (W * H) * e["dist_loss"].sum()), e["n_contrib"]

    e = _ext(case, o, np.float64, n_samp, geos, rgbs, need_depth=True, need_dist=True, need_normal=True)
    dg, dr, _ = O.render_bwd_ext(cam, o["ranges"], o["vox_list"], o["bbox"], sc["vox_centers"],
                                 sc["vox_lengths"], geos, rgbs, e["T"], e["n_contrib"], e["tile_last"], wc, wt,
                                 n_samp=n_samp, need_depth=True, need_normal=True, dL_ddepth=wd, dL_dnormal=wn,
                                 lambda_dist=lam, out_depth=e["depth"], dtype=np.float64)
    _, nc0 = loss(geos, rgbs)
    touched = np.unique(o["vox_list"] & 0x1FFFFFFF)
    eps = 1e-6
    checked, bad = 0, []
    for k in range(160):
        v = int(rng.choice(touched))
        g, r = geos.copy(), rgbs.copy()
        if k % 4 != 3:
            i = int(rng.integers(8))
            g[v, i] += eps; Lp, ncp = loss(g, r)
            g[v, i] -= 2 * eps; Lm, ncm = loss(g, r)
            ana = dg[v, i]
        else:
            ch = int(rng.integers(3))
            r[v, ch] += eps; Lp, ncp = loss(g, r)
            r[v, ch] -= 2 * eps; Lm, ncm = loss(g, r)
            ana = dr[v, ch]
        if not (np.array_equal(ncp, nc0) and np.array_equal(ncm, nc0)):
            continue
        fd = (Lp - Lm) / (2 * eps)
        checked += 1
        if abs(fd - ana) > 1e-4 * max(abs(fd), abs(ana)) + 1e-7:
            bad.append((k, v, fd, ana))
    assert checked >= 140, checked
    assert not bad, bad[:5]


def test_regularizers_closed_form():
    """R_concen adds lambda/(WH) * w * 2 (c - gt) to each colour gradient; ascending adds
    lambda/(WH) * w * [d(a) > d(b)] * (tri(a) - tri(b)) to the densities (w = entry weight)."""
    case, o = _small_case(4)
    sc, cam = case["scene"], case["cam"]
    W, H = cam["W"], cam["H"]
    rng = np.random.default_rng(9)
    gt = rng.uniform(0, 1, (3, H, W))
    zero3 = np.zeros((3, H, W))
    e = _ext(case, o, np.float64, 1, sc["geos"].astype(np.float64), o["rgbs"].astype(np.float64))
    args = (cam, o["ranges"], o["vox_list"], o["bbox"], sc["vox_centers"], sc["vox_lengths"],
            sc["geos"].astype(np.float64), o["rgbs"].astype(np.float64), e["T"], e["n_contrib"], e["tile_last"],
            zero3, np.zeros((H, W)))
    g0, r0, _ = O.render_bwd_ext(*args, dtype=np.float64)
    g1, r1, _ = O.render_bwd_ext(*args, lambda_R=0.8, gt=gt, lambda_asc=0.6, dtype=np.float64)
    pix = [(x, y) for y in range(H) for x in range(W)]
    bf, _ = BF.render_ext(cam, sc["octpath"], o["n_dup"], o["bbox"], o["qbits"], sc["vox_centers"],
                          sc["vox_lengths"], sc["geos"], o["rgbs"], pix)
    exp_r = np.zeros_like(r0)
    exp_g = np.zeros_like(g0)
    for (x, y), res in bf.items():
        ro, rd = BF.ray(x, y, cam)
        for v, w, a, b in res["entries"]:
            exp_r[v] += 0.8 / (W * H) * w * 2 * (o["rgbs"][v].astype(np.float64) - gt[:, y, x])
            c, l = sc["vox_centers"][v].astype(np.float64), float(np.asarray(sc["vox_lengths"][v]).reshape(-1)[0])
            qa = (ro + a * rd - (c - l / 2)) / l
            qb = (ro + b * rd - (c - l / 2)) / l
            tri = lambda q: np.array([(q[0] if i & 4 else 1 - q[0]) * (q[1] if i & 2 else 1 - q[1]) *  # noqa
                                      (q[2] if i & 1 else 1 - q[2]) for i in range(8)])
            g = sc["geos"][v].astype(np.float64)
            if BF.density(g, qa) > BF.density(g, qb):
                exp_g[v] += 0.6 / (W * H) * w * (tri(qa) - tri(qb))
    assert np.abs((r1 - r0) - exp_r).max() <= 1e-6 * max(1e-3, np.abs(exp_r).max())
    assert np.abs((g1 - g0) - exp_g).max() <= 1e-5 * max(1e-3, np.abs(exp_g).max())
    assert np.abs(exp_g).max() > 0 and np.abs(exp_r).max() > 0
