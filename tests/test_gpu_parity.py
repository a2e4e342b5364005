"""HIP path vs the oracle on seeded scenes (calls go through the C ABI).

Bar: integer outputs (n_duplicates, bboxes, quadrant bits, 128-bit sort keys,
sorted values, tile ranges, n_contrib, tile_last) bit-exact; images within
1e-5 relative; gradients within 1e-4 relative (elementwise, with an absolute
floor of 1e-3 x max|g| for entries near zero; the GPU sums gradients with
float atomics in a different order than the oracle's float64 sum).
"""
import numpy as np
import pytest

from tests._harness import make_case, rel_err, run_gpu, run_oracle

IMG_RTOL = 1e-5
GRAD_RTOL = 1e-4


def _check_forward(g, o):
    assert np.array_equal(g["n_dup"].astype(np.uint32), o["n_dup"])
    assert np.array_equal(g["geom_n_dup"], o["n_dup"])
    live = o["n_dup"] > 0
    assert np.array_equal(g["bbox"][live], o["bbox"][live])
    assert np.array_equal(g["qbits"][live], o["qbits"][live])
    assert g["R"] == o["vox_list"].size
    assert np.array_equal(g["idx"], o["idx"])
    assert np.array_equal(g["keys_hi"], o["keys_hi"])
    assert np.array_equal(g["keys_lo"], o["keys_lo"])
    assert np.array_equal(g["vox_list"], o["vox_list"])
    assert np.array_equal(g["ranges"], o["ranges"])
    assert np.array_equal(g["n_contrib"], o["n_contrib"])
    assert np.array_equal(g["tile_last"], o["tile_last"])
    assert rel_err(g["rgbs"], o["rgbs"], 0) <= IMG_RTOL
    assert rel_err(g["color"], o["color"], 0) <= IMG_RTOL
    assert rel_err(g["T"], o["T"], 0) <= IMG_RTOL


def _check_backward(g, o):
    for k in ("dgeos", "drgbs", "dsub", "dsh0", "dshs"):
        e = rel_err(g[k], o[k])
        assert e <= GRAD_RTOL, f"{k}: rel err {e:.3g}"


@pytest.mark.gpu
@pytest.mark.parametrize("name,seed", [("C", 0), ("C", 1), ("C", 2), ("A", 0)])
def test_parity_small(name, seed):
    case = make_case(name, seed=seed)
    print(f"case {name} seed {seed} N={case['scene']['octpath'].shape[0]}", flush=True)
    o = run_oracle(case)
    g = run_gpu(case)
    _check_forward(g, o)
    _check_backward(g, o)


@pytest.mark.gpu
@pytest.mark.parametrize("name,direction", [("C", (0.02, 0.03, 1.0)), ("C", (1.0, -0.01, 0.02)),
                                            ("A", (0.01, -0.02, -1.0))])
def test_parity_multi_quadrant(name, direction):
    """Camera close to a coordinate axis: the view straddles axis planes, so pixels
    of one tile have different ray quadrants and tile lists interleave quadrants;
    voxels cut by those planes have several quadrants (one entry run each)."""
    case = make_case(name, seed=3, direction=direction)
    o = run_oracle(case)
    assert np.unique(o["vox_list"] >> 29).size >= 4
    live_bits = o["qbits"][o["n_dup"] > 0].astype(np.uint32)
    assert (np.unpackbits(live_bits.view(np.uint8)).reshape(-1, 32).sum(1) >= 2).sum() >= 100
    g = run_gpu(case)
    _check_forward(g, o)
    _check_backward(g, o)


@pytest.mark.gpu
@pytest.mark.parametrize("n_samp", [2, 3])
def test_parity_nsamp(n_samp):
    case = make_case("C", seed=5)
    o = run_oracle(case, n_samp=n_samp)
    g = run_gpu(case, n_samp=n_samp)
    _check_forward(g, o)
    _check_backward(g, o)


@pytest.mark.gpu
def test_parity_white_background():
    case = make_case("C", seed=7)
    o = run_oracle(case, bg=1.0)
    g = run_gpu(case, bg=1.0)
    _check_forward(g, o)
    _check_backward(g, o)


@pytest.mark.gpu
def test_parity_config_b_full():
    """1M voxels (with the 5-level outside shell), 1920x1080."""
    case = make_case("B", seed=0)
    print("case B seed 0", flush=True)
    o = run_oracle(case)
    g = run_gpu(case)
    _check_forward(g, o)
    _check_backward(g, o)


@pytest.mark.gpu
def test_parity_axis_aligned_rays():
    """Camera on the z axis with an axis-aligned frame and an odd image size: the
    centre row and column rays have a zero direction component (1/rd infinite)
    and start on the octree's x = 0 and y = 0 face planes, so their ray-box
    slabs produce NaN and the reference's ternary min/max decides the hits.
    Tiles holding such rays take the exact ray-box form, the others the
    hardware min/max; both must match the oracle bit for bit."""
    from sv_raster_amd import synthetic as S
    cfg = dict(S.CONFIGS["C"], views=[(63, 63)])
    case = make_case("C", seed=4, cfg=cfg, direction=(0.0, 0.0, -1.0))
    c2w = case["cam"]["c2w"]
    assert c2w[0, 3] == 0.0 and c2w[1, 3] == 0.0 and c2w[0, 1] == 0.0 and c2w[0, 2] == 0.0
    o = run_oracle(case)
    g = run_gpu(case)
    _check_forward(g, o)
    _check_backward(g, o)


@pytest.mark.gpu
@pytest.mark.parametrize("n,direction", [(5000, None), (4096, None), (8193, None), (3000, (0.0, 0.0, -1.0))])
def test_visible_voxels(n, direction):
    """svr_visible_voxels: the in-frustum list equals torch.where(n_dup > 0)[0] and R equals svr_num_rendered,
    at voxel counts on and off the 4096-voxel scan block, including a camera that sees nothing of the scene
    (placed outside it, looking away)."""
    import torch
    from sv_raster_amd import _C, synthetic as S
    sc = S.make_scene("A", seed=4, n=n)
    cam = S.make_camera(128, 96, seed=4, direction=direction)
    if direction is not None:  # look away from the scene: every voxel behind the camera
        cam["w2c"] = np.linalg.inv(cam["c2w"].astype(np.float64) @ np.diag([-1.0, 1.0, -1.0, 1.0])).astype(np.float32)
        cam["c2w"] = np.linalg.inv(cam["w2c"].astype(np.float64)).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    centers, lengths, octpath = t(sc["vox_centers"]), t(sc["vox_lengths"]), t(sc["octpath"])
    N = centers.shape[0]
    print(f"visible_voxels N={N} direction={direction}", flush=True)
    n_dup, geom = _C.rasterize_preprocess(cam["W"], cam["H"], cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"],
                                          t(cam["w2c"]), t(cam["c2w"]), cam["near"], octpath, centers, lengths)
    idx, R = _C.visible_voxels(geom, N, centers.device)
    ref = torch.where(n_dup > 0)[0]
    assert idx.dtype == torch.int64 and torch.equal(idx, ref)
    assert R == int(n_dup.long().sum().item())
    if direction is not None:
        assert idx.numel() == 0 and R == 0


@pytest.mark.gpu
@pytest.mark.parametrize("name,seed", [("C", 5), ("A", 1)])
def test_renderer_prepared_binning_matches_oracle(name, seed):
    """The drop-in renderer.rasterize_voxels, whose visible_voxels call queues the pair offsets before the host
    synchronisation (forward called with SVR_FWD_PAIRS_READY): image, T and the density / SH gradients of one
    MSE step against the oracle, twice in a row on the same octree (cached order table)."""
    import torch
    from oracle import oracle as O
    from sv_raster_amd import renderer, synthetic as S
    case = make_case(name, seed=seed)
    sc, cam = case["scene"], case["cam"]
    print(f"renderer prepared binning {name} seed {seed}", flush=True)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    geos, sh0, shs = t(sc["geos"]).requires_grad_(), t(sc["sh0"]).requires_grad_(), t(sc["shs"]).requires_grad_()
    settings = renderer.RasterSettings(color_mode="sh", n_samp_per_vox=1, image_width=cam["W"],
                                       image_height=cam["H"], tanfovx=cam["tanfovx"], tanfovy=cam["tanfovy"],
                                       cx=cam["cx"], cy=cam["cy"], w2c_matrix=t(cam["w2c"]), c2w_matrix=t(cam["c2w"]))

    def vox_fn(idx, cam_pos, color_mode):
        rgbs = renderer.SH_eval.apply(3, idx, centers, cam_pos, None, sh0, shs)
        return {"geos": geos, "rgbs": rgbs, "subdiv_p": torch.ones((centers.shape[0], 1), device="cuda")}

    o = run_oracle(case)
    for _ in range(2):
        geos.grad = sh0.grad = shs.grad = None
        color, _, _, T, _ = renderer.rasterize_voxels(settings, octpath, centers, lengths, vox_fn)
        loss = (color * t(case["dL_dcolor"])).sum() + (T * t(case["dL_dT"])).sum()
        loss.backward()
        torch.cuda.synchronize()
        assert rel_err(color.detach().cpu().numpy(), o["color"], 0) <= IMG_RTOL
        assert rel_err(T.detach().cpu().numpy().reshape(o["T"].shape), o["T"], 0) <= IMG_RTOL
        assert rel_err(geos.grad.cpu().numpy(), o["dgeos"]) <= GRAD_RTOL
        assert rel_err(sh0.grad.cpu().numpy(), o["dsh0"]) <= GRAD_RTOL
        assert rel_err(shs.grad.cpu().numpy(), o["dshs"]) <= GRAD_RTOL


@pytest.mark.gpu
def test_backward_without_T_loss_equals_zero_T_gradient():
    """A loss on the colour only: the drop-in autograd no longer materialises a zero T gradient, the C ABI
    takes dL_dout_T = NULL; the gradients equal those of an explicit zero dL_dout_T, through
    rasterize_voxels (config A, 800x800); within the gradient bar, since the LDS float atomics may add in a
    different order from one run to the next."""
    import torch
    from oracle import oracle as O
    from sv_raster_amd import renderer, synthetic as S
    sc = S.make_scene("A", seed=2)
    cam = S.make_camera(800, 800, seed=2)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    w2c, c2w = t(cam["w2c"]), t(cam["c2w"])
    rgbs = t(O.sh_fwd(3, None, sc["vox_centers"], cam["c2w"][:3, 3], sc["sh0"], sc["shs"]).astype(np.float32))
    target = t(S.make_target(800, 800, 2))
    settings = renderer.RasterSettings(color_mode="sh", n_samp_per_vox=1, image_width=800, image_height=800,
                                       tanfovx=cam["tanfovx"], tanfovy=cam["tanfovy"], cx=cam["cx"], cy=cam["cy"],
                                       w2c_matrix=w2c, c2w_matrix=c2w)
    grads = []
    for use_T in (False, True):
        geos = t(sc["geos"]).requires_grad_()
        rg = rgbs.clone().requires_grad_()

        def vox_fn(idx, cam_pos, color_mode):
            return {"geos": geos, "rgbs": rg, "subdiv_p": torch.ones((centers.shape[0], 1), device="cuda")}

        color, _, _, T, _ = renderer.rasterize_voxels(settings, octpath, centers, lengths, vox_fn)
        loss = torch.nn.functional.mse_loss(color, target) + (0.0 * T.sum() if use_T else 0.0)
        loss.backward()
        grads.append((geos.grad.clone(), rg.grad.clone()))
    torch.cuda.synchronize()
    for a, b in zip(*grads):
        assert a.abs().max() > 0
        assert rel_err(a.cpu().numpy(), b.cpu().numpy()) <= GRAD_RTOL
