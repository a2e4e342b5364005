"""HIP compositor options (mean / median depth, distortion sum, normal, max_w)
and the optional backward terms (depth, normal, distortion, R_concen,
ascending) vs the oracle, through the C ABI.

Bar: n_contrib and tile_last bit-exact; the depth, distortion, normal and
max_w outputs bit-exact (they depend on the geometry alone, which the kernels
evaluate in the oracle's operation order); colour and T within 1e-5
relative; gradients within 1e-4 relative (float atomics vs the oracle's
float64 sums, same floor as the plain parity tests).
"""
import numpy as np
import pytest

from tests._harness import make_case, rel_err

IMG_RTOL = 1e-5
GRAD_RTOL = 1e-4
FLAGS_ALL = dict(need_depth=True, need_dist=True, need_normal=True, track_max_w=True)


def _inputs(case, seed):
    H, W = case["cam"]["H"], case["cam"]["W"]
    rng = np.random.default_rng(seed + 1234)
    f = lambda *s: (rng.normal(0, 1, s) / (H * W)).astype(np.float32)  # noqa: E731
    return dict(dL_ddepth=f(3, H, W), dL_dnormal=f(3, H, W), gt=rng.uniform(0, 1, (3, H, W)).astype(np.float32))


def _oracle(case, n_samp, fl, reg, extra, bg=0.0):
    from oracle import oracle as O
    from tests._harness import run_oracle
    sc, cam = case["scene"], case["cam"]
    o = run_oracle(case, bg=bg, n_samp=n_samp, backward=False)
    e = O.render_fwd_ext(cam, o["ranges"], o["vox_list"], o["bbox"], sc["vox_centers"], sc["vox_lengths"],
                         sc["geos"], o["rgbs"], bg=bg, n_samp=n_samp, **fl)
    res = dict(e, rgbs=o["rgbs"])
    if reg is not None:
        dg, dr, ds = O.render_bwd_ext(cam, o["ranges"], o["vox_list"], o["bbox"], sc["vox_centers"], sc["vox_lengths"],
                                      sc["geos"], o["rgbs"], e["T"], e["n_contrib"], e["tile_last"],
                                      case["dL_dcolor"], case["dL_dT"].reshape(e["T"].shape), bg=bg, n_samp=n_samp,
                                      need_depth=reg["need_depth"], need_normal=reg["need_normal"],
                                      dL_ddepth=extra["dL_ddepth"], dL_dnormal=extra["dL_dnormal"],
                                      lambda_R=reg["lambda_R"], gt=extra["gt"], lambda_asc=reg["lambda_asc"],
                                      lambda_dist=reg["lambda_dist"], out_depth=e["depth"])
        res.update(dgeos=dg, drgbs=dr, dsub=ds)
    return res


def _gpu(case, n_samp, fl, reg, extra, bg=0.0):
    import torch
    from sv_raster_amd import _C
    sc, cam = case["scene"], case["cam"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    geos, sh0, shs = t(sc["geos"]), t(sc["sh0"]), t(sc["shs"])
    w2c, c2w = t(cam["w2c"]), t(cam["c2w"])
    W, H = cam["W"], cam["H"]
    n_dup, geom = _C.rasterize_preprocess(W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w,
                                          cam["near"], octpath, centers, lengths)
    idx = torch.where(n_dup > 0)[0]
    rgbs = _C.sh_compute(3, idx, centers, c2w[:3, 3], sh0, shs)
    R, binning, img, color, depth, normal, T, max_w = _C.rasterize_voxels(
        n_samp, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, bg, fl["need_depth"],
        fl["need_dist"], fl["need_normal"], fl["track_max_w"], octpath, centers, lengths, geos, rgbs, geom)
    _, tile_last, n_contrib = _C.unpack_ImageState(W, H, img)
    out = dict(color=color, T=T.reshape(H, W), depth=depth, normal=normal, max_w=max_w.reshape(-1),
               n_contrib=n_contrib, tile_last=tile_last.reshape(-1))
    if reg is not None:
        dgeos, drgbs, dsub = _C.rasterize_voxels_backward(
            R, n_samp, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, bg, octpath, centers,
            lengths, geos, rgbs, geom, binning, img, T, t(case["dL_dcolor"]), t(extra["dL_ddepth"]),
            t(extra["dL_dnormal"]), t(case["dL_dT"]), reg["lambda_R"], t(extra["gt"]), reg["lambda_asc"],
            reg["lambda_dist"], reg["need_depth"], reg["need_normal"], depth, normal)
        out.update(dgeos=dgeos, drgbs=drgbs, dsub=dsub)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    for k in ("n_contrib", "tile_last"):
        res[k] = res[k].view(np.uint32)
    return res


def _check_fwd(g, o, fl):
    assert np.array_equal(g["n_contrib"], o["n_contrib"])
    assert np.array_equal(g["tile_last"], o["tile_last"])
    assert rel_err(g["color"], o["color"], 0) <= IMG_RTOL
    assert rel_err(g["T"], o["T"], 0) <= IMG_RTOL
    if fl["need_depth"] or fl["need_dist"]:
        for ch, on in ((0, fl["need_depth"]), (1, fl["need_dist"]), (2, fl["need_depth"])):
            ref = o["depth"][ch] if on else np.zeros_like(o["depth"][ch])
            bad = np.count_nonzero(g["depth"][ch] != ref)
            assert bad == 0, f"depth channel {ch}: {bad} pixels differ, max {np.abs(g['depth'][ch] - ref).max():.3g}"
    else:
        assert g["depth"].size == 0
    if fl["need_normal"]:
        bad = np.count_nonzero(g["normal"] != o["normal"])
        assert bad == 0, f"normal: {bad} differ, max {np.abs(g['normal'] - o['normal']).max():.3g}"
    else:
        assert g["normal"].size == 0
    if fl["track_max_w"]:
        assert np.array_equal(g["max_w"], o["max_w"]), np.abs(g["max_w"] - o["max_w"]).max()
    else:
        assert g["max_w"].size == 0


def _check_bwd(g, o):
    for k in ("dgeos", "drgbs", "dsub"):
        e = rel_err(g[k], np.asarray(o[k]).reshape(g[k].shape))
        assert e <= GRAD_RTOL, f"{k}: rel err {e:.3g}"


REG_ALL = dict(need_depth=True, need_normal=True, lambda_R=0.5, lambda_asc=0.3, lambda_dist=0.2)


@pytest.mark.gpu
@pytest.mark.parametrize("name,seed,n_samp", [("C", 0, 1), ("C", 1, 2), ("C", 2, 3), ("A", 0, 1)])
def test_ext_all_options(name, seed, n_samp):
    case = make_case(name, seed=seed)
    print(f"ext all options {name} seed {seed} n_samp {n_samp}", flush=True)
    extra = _inputs(case, seed)
    o = _oracle(case, n_samp, FLAGS_ALL, REG_ALL, extra)
    g = _gpu(case, n_samp, FLAGS_ALL, REG_ALL, extra)
    _check_fwd(g, o, FLAGS_ALL)
    _check_bwd(g, o)


@pytest.mark.gpu
@pytest.mark.parametrize("opt", ["need_depth", "need_dist", "need_normal", "track_max_w"])
def test_ext_single_forward_option(opt):
    """Each forward option alone: its outputs match, the others stay empty / zero."""
    case = make_case("C", seed=3)
    fl = {k: k == opt for k in FLAGS_ALL}
    print(f"ext forward option {opt}", flush=True)
    o = _oracle(case, 1, fl, None, None)
    g = _gpu(case, 1, fl, None, None)
    _check_fwd(g, o, fl)


@pytest.mark.gpu
@pytest.mark.parametrize("term", ["need_depth", "need_normal", "lambda_R", "lambda_asc", "lambda_dist"])
def test_ext_single_backward_term(term):
    """Each backward term alone on top of the colour / T gradient."""
    case = make_case("C", seed=4)
    extra = _inputs(case, 4)
    reg = {k: (REG_ALL[k] if k == term else (False if k.startswith("need") else 0.0)) for k in REG_ALL}
    fl = dict(FLAGS_ALL, track_max_w=False)
    print(f"ext backward term {term}", flush=True)
    o = _oracle(case, 1, fl, reg, extra)
    g = _gpu(case, 1, fl, reg, extra)
    _check_bwd(g, o)


@pytest.mark.gpu
def test_ext_white_background_axis_aligned():
    """bg = 1 and an axis-aligned view (rays with a zero direction component take the exact ray-box launch)."""
    case = make_case("C", seed=6, direction=(0.0, 0.0, -1.0))
    extra = _inputs(case, 6)
    print("ext bg 1 axis-aligned", flush=True)
    o = _oracle(case, 1, FLAGS_ALL, REG_ALL, extra, bg=1.0)
    g = _gpu(case, 1, FLAGS_ALL, REG_ALL, extra, bg=1.0)
    _check_fwd(g, o, FLAGS_ALL)
    _check_bwd(g, o)


@pytest.mark.gpu
def test_ext_autograd_through_renderer():
    """renderer.rasterize_voxels with depth / normal outputs and the regularizers: autograd's gradients of
    <wc, color> + <wd, depth> + <wn, normal> + <wt, T> equal the oracle's adjoint."""
    import torch
    from sv_raster_amd import renderer
    from tests.test_gpu_batch import _settings
    case = make_case("C", seed=8)
    sc, cam = case["scene"], case["cam"]
    H, W = cam["H"], cam["W"]
    extra = _inputs(case, 8)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    o = _oracle(case, 1, dict(FLAGS_ALL, track_max_w=False), REG_ALL, extra)
    settings = _settings(cam)._replace(need_depth=True, need_normal=True, track_max_w=True,
                                       lambda_R_concen=REG_ALL["lambda_R"], lambda_ascending=REG_ALL["lambda_asc"],
                                       lambda_dist=REG_ALL["lambda_dist"], gt_color=t(extra["gt"]))
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    geos = t(sc["geos"]).requires_grad_()
    rgbs_full = np.zeros((centers.shape[0], 3), np.float32)
    rgbs_full[:] = o["rgbs"]
    rgbs = t(rgbs_full).requires_grad_()

    def vox_fn(idx, cam_pos, color_mode):
        return {"geos": geos, "rgbs": rgbs, "subdiv_p": torch.ones((centers.shape[0], 1), device="cuda")}

    color, depth, normal, T, max_w = renderer.rasterize_voxels(settings, octpath, centers, lengths, vox_fn)
    assert depth.shape == (3, H, W) and normal.shape == (3, H, W) and max_w.shape == (centers.shape[0], 1)
    loss = ((color * t(case["dL_dcolor"])).sum() + (depth * t(extra["dL_ddepth"])).sum() +
            (normal * t(extra["dL_dnormal"])).sum() + (T * t(case["dL_dT"])).sum())
    loss.backward()
    torch.cuda.synchronize()
    for got, ref, name in ((geos.grad, o["dgeos"], "geos"), (rgbs.grad, o["drgbs"], "rgbs")):
        e = rel_err(got.cpu().numpy(), np.asarray(ref))
        assert e <= GRAD_RTOL, f"{name}: rel err {e:.3g}"
    assert np.array_equal(depth.detach().cpu().numpy()[0], o["depth"][0])
