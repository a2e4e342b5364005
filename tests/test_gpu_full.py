"""The multi-view batch at full size against the single-view drop-in path.

At the sizes of the multi-view workloads the oracle would take hours, so these
tests check a size-independent property instead: the batch (one launch per
stage over all views) gives every view the image, transmittance, per-pixel
contributor counts and tile ranges of the single-view path run on that view
alone, and gradients equal to the sum over views of the single-view gradients.
The single-view path is itself held to the oracle up to the 1 M-voxel 1080p
scene (test_gpu_parity.py).  Workloads: the 2 M-voxel scene with its 16 views
of uneven resolution, and the 8 M-voxel scene with the outside shell and all
64 of its views in one batch.

Bar: n_contrib and tile ranges bit-exact, images within 1e-5 relative,
gradients within 1e-4 (elementwise, absolute floor 1e-3 x max|g|)."""
import numpy as np
import pytest

from sv_raster_amd import synthetic as S

IMG_RTOL = 1e-5
GRAD_RTOL = 1e-4


def _rel_err(a, b, floor_frac=1e-3):
    """max |a - b| / (|b| + floor_frac * max|b|), on the device."""
    import torch
    a = a.double()
    b = b.double()
    scale = b.abs().max()
    if float(scale) == 0.0:
        return float(a.abs().max()) if a.numel() else 0.0
    diff = (a - b).abs()
    r = torch.where(diff == 0, torch.zeros_like(diff), diff / (b.abs() + floor_frac * scale))
    return float(r.max())


def _full_case(name, n_views):
    import torch
    sc = S.make_scene(name, seed=0)
    sizes = S.config_views(name, 0)[:n_views]
    cams = [S.make_camera(W, H, seed=1000 + u) for u, (W, H) in enumerate(sizes)]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dev = dict(octpath=t(sc["octpath"]), centers=t(sc["vox_centers"]), lengths=t(sc["vox_lengths"]),
               geos=t(sc["geos"]), sh0=t(sc["sh0"]), shs=t(sc["shs"]))
    del sc
    g = torch.Generator(device="cuda").manual_seed(7)
    dcol = [torch.randn((3, c["H"], c["W"]), device="cuda", generator=g) / (c["H"] * c["W"]) for c in cams]
    dT = [torch.randn((1, c["H"], c["W"]), device="cuda", generator=g) / (c["H"] * c["W"]) for c in cams]
    return dev, cams, dcol, dT


def _settings(cam):
    import torch
    from sv_raster_amd import renderer
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    return renderer.RasterSettings(color_mode="sh", n_samp_per_vox=1, image_width=cam["W"], image_height=cam["H"],
                                   tanfovx=cam["tanfovx"], tanfovy=cam["tanfovy"], cx=cam["cx"], cy=cam["cy"],
                                   w2c_matrix=t(cam["w2c"]), c2w_matrix=t(cam["c2w"]), near=cam["near"])


def _single_view(dev, cam, dcol, dT):
    """The drop-in _C path on one view: preprocess, visible list, SH, forward, backward, SH backward."""
    import torch
    from sv_raster_amd import _C
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    W, H = cam["W"], cam["H"]
    w2c, c2w = t(cam["w2c"]), t(cam["c2w"])
    N = dev["centers"].shape[0]
    n_dup, geom = _C.rasterize_preprocess(W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w,
                                          cam["near"], dev["octpath"], dev["centers"], dev["lengths"])
    table = _C.order_table_cache.get(dev["octpath"])
    idx, R = _C.visible_voxels(geom, N, dev["centers"].device, table)
    cam_pos = c2w[:3, 3]
    rgbs = _C.sh_compute(3, idx, dev["centers"], cam_pos, dev["sh0"], dev["shs"])
    R, binning, img, color, _, _, T, _ = _C.rasterize_voxels(
        1, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, 0.0, False, False, False, False,
        dev["octpath"], dev["centers"], dev["lengths"], dev["geos"], rgbs, geom, num_rendered=R, pairs_ready=True)
    ranges, tile_last, n_contrib = _C.unpack_ImageState(W, H, img)
    dgeos, drgbs, dsub = _C.rasterize_voxels_backward(
        R, 1, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, 0.0, dev["octpath"],
        dev["centers"], dev["lengths"], dev["geos"], rgbs, geom, binning, img, T, dcol, None, None, dT, 0.0, None,
        0.0, 0.0, False, False, None, None)
    dsh0, dshs = _C.sh_compute_bw(3, 16, idx, dev["centers"], cam_pos, rgbs, drgbs)
    return dict(R=R, n_dup=n_dup, color=color, T=T, rgbs=rgbs, drgbs=drgbs, ranges=ranges.reshape(-1, 2),
                tile_last=tile_last.reshape(-1), n_contrib=n_contrib.reshape(-1), dgeos=dgeos, dsub=dsub,
                dsh0=dsh0, dshs=dshs)


def _check(name, n_views, tile_shape="16x16", case=None, bwd=(4, 8)):
    """Batch against the single-view path, view by view.  With the reference's 16x16 binning tile the batch's
    tile ranges and per-pixel n_contrib (positions in a tile's list) are compared bit for bit; with the 8x8
    and 32x8 tiles the lists are those of other tiles, so the per-pixel outputs (image, T) and the gradients
    carry the comparison."""
    import torch
    from sv_raster_amd import batch
    dev, cams, dcol, dT = _full_case(name, n_views) if case is None else case
    N = dev["centers"].shape[0]
    print(f"full {name}: {N} voxels, {n_views} views, {tile_shape} tiles", flush=True)
    views = batch.ViewBatch([_settings(c) for c in cams], tile_shape=tile_shape, bwd_window=bwd[0], bwd_sync=bwd[1])
    same_tiles = tile_shape == "16x16"
    st = batch.batch_forward(views, 3, dev["octpath"], dev["centers"], dev["lengths"], dev["geos"], dev["sh0"],
                             dev["shs"])
    bdg, bdsh0, bdshs, bdsub, bdrgbs = batch.batch_backward(views, st, torch.cat([d.reshape(-1) for d in dcol]),
                                                            torch.cat([d.reshape(-1) for d in dT]))
    b_ranges, b_tile_last, b_n_contrib = batch.unpack_image_state(views, st["img"])
    sums = {k: None for k in ("dgeos", "dsh0", "dshs", "dsub")}
    r_off, t_off = 0, 0
    for u, cam in enumerate(cams):
        W, H = cam["W"], cam["H"]
        P = W * H
        pb = views.pix_base[u]
        T_u = ((W + 15) // 16) * ((H + 15) // 16)
        s = _single_view(dev, cam, dcol[u], dT[u])
        print(f"  view {u}: {W}x{H}, R {s['R']}", flush=True)
        assert _rel_err(st["color"][3 * pb:3 * pb + 3 * P], s["color"].reshape(-1), 0) <= IMG_RTOL, u
        assert _rel_err(st["T"][pb:pb + P], s["T"].reshape(-1), 0) <= IMG_RTOL, u
        if same_tiles:
            assert torch.equal(st["n_dup"][u * N:(u + 1) * N].int(), s["n_dup"].int()), u
        else:  # tile counts differ with the tile, the set of voxels with entries does not
            assert torch.equal(st["n_dup"][u * N:(u + 1) * N] > 0, s["n_dup"] > 0), u
        if same_tiles:
            assert torch.equal(b_n_contrib[pb:pb + P].int(), s["n_contrib"].int()), u
            rg = b_ranges[t_off:t_off + T_u].long()
            ro = s["ranges"].long()
            nonempty = ro[:, 1] > ro[:, 0]
            assert torch.equal(rg[nonempty], ro[nonempty] + r_off), u
            assert torch.equal(b_tile_last[t_off:t_off + T_u].long()[nonempty],
                               s["tile_last"].long()[nonempty] + r_off), u
        assert _rel_err(bdrgbs[u * N:(u + 1) * N], s["drgbs"]) <= GRAD_RTOL, u
        for k in sums:
            v = s[k].double()
            sums[k] = v if sums[k] is None else sums[k] + v
        r_off += s["R"]
        t_off += T_u
        del s
    if same_tiles:
        assert st["R"] == r_off
    for k, g in (("dgeos", bdg), ("dsh0", bdsh0), ("dshs", bdshs), ("dsub", bdsub)):
        e = _rel_err(g.reshape(sums[k].shape), sums[k])
        assert e <= GRAD_RTOL, f"{k}: rel err {e:.3g}"


@pytest.fixture(scope="module")
def case_d():
    return _full_case("D", 16)


@pytest.mark.gpu
@pytest.mark.parametrize("tile_shape", ["16x16", "32x8", "8x8"])
def test_batch_matches_single_view_full_multiview_2m(case_d, tile_shape):
    """2 M voxels (levels 4-14), 16 views of 640x480 ... 2560x1440 (the multi-view bench workload), at every
    binning tile the autotuner can pick."""
    _check("D", 16, tile_shape, case=case_d)


@pytest.mark.gpu
@pytest.mark.parametrize("tile_shape", ["32x8", "16x16"])
def test_graph_replays_match_eager_full_multiview_2m(case_d, tile_shape):
    """The captured training step (GraphTrainStep) at the multi-view bench workload, three replays.  Before
    each replay the eager batch path (rasterize_voxels_batch, summed MSE, autograd) runs from the same
    parameters: the replay's images are within 1e-5 and its gradients within 1e-4 of the eager ones.  The
    replay's Adam update is then checked bit for bit against the oracle's Adam (adam_step.cu restated) applied
    on the host to the replay's own gradients, parameters and moments carried along on the host, so the
    parameters never drift from an exact reference.  (Comparing parameters after Adam with an independent eager
    run is not a meaningful bar: with eps 1e-15 Adam scales gradients that nearly cancel by 1 / |g|, which
    turns their summation-order rounding into visible parameter differences.)"""
    import torch
    from oracle import oracle as O
    from sv_raster_amd import batch, sparse_adam
    dev, cams, _, _ = case_d
    settings = [_settings(c) for c in cams]
    views = batch.ViewBatch(settings, tile_shape=tile_shape)
    targets = [torch.from_numpy(S.make_target(c["H"], c["W"], u)).cuda() for u, c in enumerate(cams)]
    pg = [dev["geos"].clone().requires_grad_(), dev["sh0"].clone().requires_grad_(),
          dev["shs"].clone().requires_grad_()]
    lrs = (0.025, 0.01, 0.01)
    opt_g = sparse_adam.SparseAdam([{"params": [p], "lr": lr} for p, lr in zip(pg, lrs)], betas=(0.1, 0.99),
                                   eps=1e-15, sparse=True)
    host = [p.detach().cpu().numpy().copy() for p in pg]
    moments = [(np.zeros_like(h), np.zeros_like(h)) for h in host]
    trainer = batch.GraphTrainStep(views, dev["octpath"], dev["centers"], dev["lengths"], pg[0], pg[1], pg[2], opt_g,
                                   targets)
    print(f"graph D {tile_shape}: capacity {trainer.capacity}", flush=True)
    for k in range(1, 4):
        pe = [p.detach().clone().requires_grad_() for p in pg]
        colors, _ = batch.rasterize_voxels_batch(views, dev["octpath"], dev["centers"], dev["lengths"], *pe)
        loss = sum(torch.nn.functional.mse_loss(c, tg) for c, tg in zip(colors, targets))
        loss.backward()
        eager_color = torch.cat([c.detach().reshape(-1) for c in colors])
        trainer.step()
        torch.cuda.synchronize()
        e = _rel_err(trainer.color, eager_color, 0)
        assert e <= IMG_RTOL, f"replay {k}: image rel err {e:.3g}"
        for g, q, name in zip(trainer.grads, pe, ("geos", "sh0", "shs")):
            e = _rel_err(g, q.grad)
            assert e <= GRAD_RTOL, f"replay {k}: d{name} rel err {e:.3g}"
        for i, (p, g) in enumerate(zip(pg, trainer.grads)):
            m, v = moments[i]
            O.adam_step(host[i], g.detach().cpu().numpy().copy(), m, v, k, lrs[i], 0.1, 0.99, 1e-15, sparse=True)
            assert np.array_equal(p.detach().cpu().numpy().view(np.uint32), host[i].view(np.uint32)), (k, i)
        print(f"  replay {k} ok", flush=True)
    assert not trainer.overflowed()


@pytest.fixture(scope="module")
def case_e():
    import torch
    case = _full_case("E", 64)
    yield case
    del case
    torch.cuda.empty_cache()


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_batch_matches_single_view_full_8m_64_views(case_e):
    """The large batched workload at full size on one GPU: 8 M voxels (levels 4-16 with the 5-level outside
    shell) and all 64 views of 720p-1440p (144 Mpix, about 590 M list entries) in one batch, against the
    single-view path view by view, with the reference's 16x16 tiles (lists and n_contrib bit-exact)."""
    import torch
    _check("E", 64, "16x16", case=case_e)
    print(f"peak device memory {torch.cuda.max_memory_allocated() / 2**30:.1f} GiB", flush=True)
    torch.cuda.empty_cache()


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_autotuned_batch_matches_single_view_full_8m_64_views(case_e):
    """The same workload in the configuration the autotuner picks for it (binning tile, backward window depth
    and sync threshold): images, T and gradients against the single-view path."""
    import torch
    from sv_raster_amd import batch
    dev, cams, _, _ = case_e
    views, rep = batch.autotune([_settings(c) for c in cams], dev["octpath"], dev["centers"], dev["lengths"],
                                dev["geos"], dev["sh0"], dev["shs"], min_rounds=2, max_rounds=4)
    print(f"autotuned: {views.tile_shape}, window {views.bwd_window}, sync {views.bwd_sync}", flush=True)
    del views
    torch.cuda.empty_cache()
    _check("E", 64, rep["tile_shape"], case=case_e, bwd=(rep["bwd_window"], rep["bwd_sync"]))
    torch.cuda.empty_cache()
