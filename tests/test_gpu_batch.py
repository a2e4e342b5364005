"""Multi-view batch path vs the oracle run view by view (calls go through the
C ABI): per view the images, transmittance and n_contrib are bit-exact /
within 1e-5, each view's tile ranges equal the single-view ranges shifted by
the entries of the previous views, and the gradients of the shared parameters
equal the sums over views of the per-view gradients within 1e-4."""
import numpy as np
import pytest

from sv_raster_amd import synthetic as S
from tests._harness import make_case, rel_err, run_oracle

IMG_RTOL = 1e-5
GRAD_RTOL = 1e-4


def _views(name, specs, seed=0):
    """One case per view ((W, H), direction or None) on the same seeded scene."""
    sc = make_case(name, seed=seed)["scene"]
    cases = []
    for u, ((W, H), direction) in enumerate(specs):
        cam = S.make_camera(W, H, seed=seed + 17 * (u + 1), direction=direction)
        rng = np.random.default_rng(seed + 99 + u)
        cases.append(dict(scene=sc, cam=cam,
                          dL_dcolor=(rng.normal(0, 1, (3, H, W)) / (H * W)).astype(np.float32),
                          dL_dT=(rng.normal(0, 1, (1, H, W)) / (H * W)).astype(np.float32)))
    return cases


def _settings(cam, bg=0.0, n_samp=1):
    import torch
    from sv_raster_amd import renderer
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    return renderer.RasterSettings(color_mode="sh", n_samp_per_vox=n_samp, image_width=cam["W"],
                                   image_height=cam["H"], tanfovx=cam["tanfovx"], tanfovy=cam["tanfovy"],
                                   cx=cam["cx"], cy=cam["cy"], w2c_matrix=t(cam["w2c"]), c2w_matrix=t(cam["c2w"]),
                                   bg_color=bg, near=cam["near"])


def _run_batch(cases, bg=0.0, n_samp=1, tile_shape="16x16", bwd=(4, 8)):
    import torch
    from sv_raster_amd import batch
    sc = cases[0]["scene"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    views = batch.ViewBatch([_settings(c["cam"], bg, n_samp) for c in cases], tile_shape=tile_shape,
                            bwd_window=bwd[0], bwd_sync=bwd[1])
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    geos, sh0, shs = t(sc["geos"]), t(sc["sh0"]), t(sc["shs"])
    st = batch.batch_forward(views, 3, octpath, centers, lengths, geos, sh0, shs)
    dcol = torch.cat([t(c["dL_dcolor"]).reshape(-1) for c in cases])
    dT = torch.cat([t(c["dL_dT"]).reshape(-1) for c in cases])
    dgeos, dsh0, dshs, dsub, drgbs = batch.batch_backward(views, st, dcol, dT)
    ranges, tile_last, n_contrib = batch.unpack_image_state(views, st["img"])
    torch.cuda.synchronize()
    out = dict(color=st["color"].cpu().numpy(), T=st["T"].cpu().numpy(), R=st["R"],
               rgbs=st["rgbs"].cpu().numpy(), n_dup=st["n_dup"].cpu().numpy(),
               ranges=ranges.cpu().numpy().view(np.uint32), tile_last=tile_last.cpu().numpy().view(np.uint32),
               n_contrib=n_contrib.cpu().numpy().view(np.uint32), dgeos=dgeos.cpu().numpy(),
               dsh0=dsh0.cpu().numpy(), dshs=dshs.cpu().numpy(), dsub=dsub.cpu().numpy(),
               drgbs=drgbs.cpu().numpy())
    return views, out


def _compare(cases, views, g, orc, tile16=True):
    """tile16: the batch bins with the reference's 16x16 tiles, so n_dup, n_contrib and the tile ranges are
    compared too (they are relative to the binning tile)."""
    N = cases[0]["scene"]["octpath"].shape[0]
    r_off = 0
    t_off = 0
    for u, (c, o) in enumerate(zip(cases, orc)):
        W, H = c["cam"]["W"], c["cam"]["H"]
        P = W * H
        pb = views.pix_base[u]
        gx, gy = (W + 15) // 16, (H + 15) // 16
        T_u = gx * gy
        col = g["color"][3 * pb:3 * pb + 3 * P].reshape(3, H, W)
        assert rel_err(col, o["color"], 0) <= IMG_RTOL, u
        assert rel_err(g["T"][pb:pb + P].reshape(H, W), o["T"], 0) <= IMG_RTOL, u
        assert rel_err(g["rgbs"][u * N:(u + 1) * N][o["idx"]], o["rgbs"][o["idx"]], 0) <= IMG_RTOL, u
        assert rel_err(g["drgbs"][u * N:(u + 1) * N], o["drgbs"]) <= GRAD_RTOL, u
        assert np.array_equal(g["n_dup"][u * N:(u + 1) * N] > 0, o["n_dup"] > 0), u
        if not tile16:
            continue
        assert np.array_equal(g["n_dup"][u * N:(u + 1) * N].astype(np.uint32), o["n_dup"]), u
        assert np.array_equal(g["n_contrib"][pb:pb + P].reshape(o["n_contrib"].shape), o["n_contrib"]), u
        rg = g["ranges"][t_off:t_off + T_u].astype(np.int64)
        ro = o["ranges"].astype(np.int64)
        nonempty = ro[:, 1] > ro[:, 0]
        assert np.array_equal(rg[nonempty], ro[nonempty] + r_off), u
        assert np.all(rg[~nonempty, 0] == rg[~nonempty, 1]), u
        tl = g["tile_last"][t_off:t_off + T_u].astype(np.int64)
        assert np.array_equal(tl[nonempty], o["tile_last"].astype(np.int64)[nonempty] + r_off), u
        r_off += o["vox_list"].size
        t_off += T_u
    if tile16:
        assert g["R"] == r_off
    for k in ("dgeos", "dsh0", "dshs", "dsub"):
        ref = np.sum([np.asarray(o[k], np.float64).reshape(g[k].shape) for o in orc], axis=0)
        e = rel_err(g[k], ref)
        assert e <= GRAD_RTOL, f"{k}: rel err {e:.3g}"


@pytest.mark.gpu
def test_batch_matches_per_view_oracle_small():
    """Four views of one config-C scene, uneven sizes (not multiples of 16), one of
    them axis-aligned so that its centre rays take the exact ray-box launch."""
    cases = _views("C", [((64, 64), None), ((48, 40), None), ((80, 56), None), ((63, 63), (0.0, 0.0, -1.0))],
                   seed=2)
    print("batch C x4", flush=True)
    orc = [run_oracle(c) for c in cases]
    views, g = _run_batch(cases)
    _compare(cases, views, g, orc)


@pytest.mark.gpu
def test_batch_matches_per_view_oracle_config_a():
    """Two config-A views (800x800 and 640x480): many tiles per view, large voxels
    (queued for whole-workgroup duplication and reduction)."""
    cases = _views("A", [((800, 800), None), ((640, 480), None)], seed=1)
    print("batch A x2", flush=True)
    orc = [run_oracle(c) for c in cases]
    views, g = _run_batch(cases)
    _compare(cases, views, g, orc)


@pytest.mark.gpu
def test_batch_autograd_matches_plain_loop():
    """rasterize_voxels_batch + autograd equals the drop-in renderer called once
    per view (SH through vox_fn), gradients summed by autograd."""
    import torch
    from sv_raster_amd import batch, renderer
    cases = _views("C", [((64, 64), None), ((32, 48), None), ((72, 40), None)], seed=3)
    sc = cases[0]["scene"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    rng = np.random.default_rng(5)
    wts = [t(rng.normal(0, 1, (3, c["cam"]["H"], c["cam"]["W"])).astype(np.float32)) for c in cases]

    def params():
        return [t(sc["geos"]).requires_grad_(), t(sc["sh0"]).requires_grad_(), t(sc["shs"]).requires_grad_()]

    geos, sh0, shs = params()
    views = batch.ViewBatch([_settings(c["cam"]) for c in cases])
    colors, Ts = batch.rasterize_voxels_batch(views, octpath, centers, lengths, geos, sh0, shs)
    loss = sum((col * w).sum() + T.sum() for col, T, w in zip(colors, Ts, wts))
    loss.backward()

    geos2, sh02, shs2 = params()
    loss2 = 0.0
    for c, w in zip(cases, wts):
        def vox_fn(idx, cam_pos, color_mode):
            rgbs = renderer.SH_eval.apply(3, idx, centers, cam_pos, None, sh02, shs2)
            return {"geos": geos2, "rgbs": rgbs, "subdiv_p": torch.ones((centers.shape[0], 1), device="cuda")}
        col, _, _, T, _ = renderer.rasterize_voxels(_settings(c["cam"]), octpath, centers, lengths, vox_fn)
        loss2 = loss2 + (col * w).sum() + T.sum()
    loss2.backward()
    for a, b in ((geos, geos2), (sh0, sh02), (shs, shs2)):
        e = rel_err(a.grad.cpu().numpy(), b.grad.cpu().numpy())
        assert e <= GRAD_RTOL, e
    l1, l2 = float(loss.detach()), float(loss2.detach())
    assert abs(l1 - l2) <= 1e-5 * max(1.0, abs(l2))


@pytest.mark.gpu
@pytest.mark.parametrize("tile_shape", ["8x8", "32x8"])
def test_batch_other_tile_shapes_match_oracle(tile_shape):
    """Binning at 8x8 or 32x8 changes the lists, not the per-pixel results: images, T and gradients still
    match the oracle (which bins at 16x16)."""
    cases = _views("C", [((64, 64), None), ((48, 40), None), ((80, 56), None), ((63, 63), (0.0, 0.0, -1.0))],
                   seed=6)
    print("batch C x4", tile_shape, flush=True)
    orc = [run_oracle(c) for c in cases]
    views, g = _run_batch(cases, tile_shape=tile_shape)
    assert views.tile_shape == tile_shape
    _compare(cases, views, g, orc, tile16=False)


@pytest.mark.gpu
@pytest.mark.parametrize("tile_shape", ["16x16", "8x8", "32x8"])
@pytest.mark.parametrize("window", [2, 4, 6, 8])
@pytest.mark.parametrize("sync", [4, 8, 16])
def test_batch_backward_depths_match_oracle(tile_shape, window, sync):
    """Every backward window depth and sync threshold the autotuner can pick gives the oracle's gradients
    (images and the forward are unaffected), on four views of the 2k-voxel scene at every binning tile."""
    cases = _views("C", [((64, 64), None), ((48, 40), None), ((80, 56), None), ((63, 63), (0.0, 0.0, -1.0))],
                   seed=6)
    print("batch C x4", tile_shape, window, sync, flush=True)
    orc = _oracle_cache("C6", cases)
    views, g = _run_batch(cases, tile_shape=tile_shape, bwd=(window, sync))
    assert (views.bwd_window, views.bwd_sync) == (window, sync)
    _compare(cases, views, g, orc, tile16=tile_shape == "16x16")


_ORACLE = {}


def _oracle_cache(key, cases):
    if key not in _ORACLE:
        _ORACLE[key] = [run_oracle(c) for c in cases]
    return _ORACLE[key]


@pytest.mark.gpu
def test_batch_backward_depths_long_lists_match_oracle():
    """Deep windows on long lists: config A (100 k voxels) seen by two 200 x 160 views, every window depth at
    the default sync threshold, against the oracle."""
    cases = _views("A", [((200, 160), None), ((160, 200), None)], seed=2)
    orc = _oracle_cache("A2", cases)
    for window in (2, 4, 6, 8):
        print("batch A x2 window", window, flush=True)
        views, g = _run_batch(cases, tile_shape="16x16", bwd=(window, 8))
        _compare(cases, views, g, orc, tile16=True)


@pytest.mark.gpu
def test_autotune_picks_and_caches():
    import torch
    from sv_raster_amd import batch
    cases = _views("C", [((64, 64), None), ((96, 48), None)], seed=7)
    sc = cases[0]["scene"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    args = (t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"]), t(sc["geos"]), t(sc["sh0"]),
            t(sc["shs"]))
    settings = [_settings(c["cam"]) for c in cases]
    views, rep = batch.autotune(settings, *args, min_rounds=2, max_rounds=3)
    assert set(rep["fwd_bwd_ms"]) == set(batch.TILE_SHAPES)
    assert len(rep["bwd_ms"]) == len(batch.BWD_WINDOWS) * len(batch.BWD_SYNCS)
    med = {k: v["median_ms"] for k, v in rep["fwd_bwd_ms"].items()}
    assert views.tile_shape == min(med, key=med.get) == rep["tile_shape"]
    bmed = {k: v["median_ms"] for k, v in rep["bwd_ms"].items()}
    assert f"{views.bwd_window}x{views.bwd_sync}" == min(bmed, key=bmed.get)
    for v in list(rep["fwd_bwd_ms"].values()) + list(rep["bwd_ms"].values()):
        assert 2 <= v["runs"] <= 3 and v["min_ms"] <= v["median_ms"] <= v["max_ms"]
    views2, rep2 = batch.autotune(settings, *args, min_rounds=2, max_rounds=3)
    assert (views2.tile_shape, views2.bwd_window, views2.bwd_sync) == (views.tile_shape, views.bwd_window,
                                                                       views.bwd_sync)
    assert rep2 == rep
