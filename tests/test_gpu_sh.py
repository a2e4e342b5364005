"""SH colour and its backward on the GPU (through the C ABI) against the oracle
(sh_compute.cu restated), for index lists that exercise every staging path of
the degree-3 kernels: runs of 64 consecutive ids at aligned and misaligned
starts (16-byte span copies), runs ending at the last voxel, gaps, unsorted
ids and duplicates (row gathers), lists of most voxels (the dense pass over
marked rows, sorted and unsorted with duplicates), and all voxels (no index
list)."""
import numpy as np
import pytest

RTOL = 1e-5


def _idx_cases(N, rng):
    perm = rng.permutation(N)
    dup = np.concatenate([np.arange(0, 40), np.arange(30, 100)])
    return {
        "run_aligned": np.arange(128, 128 + 640),
        "run_misaligned": np.arange(5, 5 + 64 * 7 + 3),
        "run_to_end": np.arange(N - 200, N),
        "gaps": np.sort(rng.choice(N, N // 3, replace=False)),
        "unsorted": perm[: N // 2],
        "duplicates": dup,
        # lists covering at least half of the voxels take the dense path (listed rows marked)
        "most_sorted": np.sort(rng.choice(N, (4 * N) // 5, replace=False)),
        "most_unsorted_duplicates": np.concatenate([perm[: (3 * N) // 4], perm[:50]]),
        "all": None,
    }


@pytest.mark.gpu
@pytest.mark.parametrize("deg", [0, 1, 3])
def test_sh_forward_backward_match_oracle(deg):
    import torch
    from oracle import oracle as O
    from sv_raster_amd import _C
    rng = np.random.default_rng(11 + deg)
    N = 3001
    centers = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    sh0 = rng.uniform(0, 1, (N, 3)).astype(np.float32)
    shs = rng.normal(0, 0.05, (N, 15, 3)).astype(np.float32)
    cam_pos = np.array([0.3, -2.5, 1.7], np.float32)
    dL = rng.normal(0, 1, (N, 3)).astype(np.float32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    for name, idx in _idx_cases(N, rng).items():
        print(f"sh deg {deg} idx {name}", flush=True)
        ti = t(idx.astype(np.int64)) if idx is not None else None
        rgbs = _C.sh_compute(deg, ti, t(centers), t(cam_pos), t(sh0), t(shs))
        ref = O.sh_fwd(deg, idx, centers, cam_pos, sh0, shs)
        got = rgbs.cpu().numpy()
        assert np.abs(got - ref).max() <= RTOL * max(1.0, np.abs(ref).max()), name
        d0, ds = _C.sh_compute_bw(deg, 16, ti, t(centers), t(cam_pos), rgbs, t(dL))
        r0, rs = O.sh_bwd(deg, 16, idx, centers, cam_pos, dL)
        assert np.abs(d0.cpu().numpy() - r0).max() <= RTOL * max(1.0, np.abs(r0).max()), name
        assert np.abs(ds.cpu().numpy() - rs.reshape(ds.shape)).max() <= RTOL * max(1.0, np.abs(rs).max()), name


@pytest.mark.gpu
@pytest.mark.parametrize("deg", [1, 3])
def test_sh_in_frustum_list_from_geometry_state(deg):
    """The SH passes over a geometry state's in-frustum voxels (svr_sh_compute_visible / _bw_visible: rows with
    n_dup > 0) equal the calls with the list svr_visible_voxels returns for that state, bit for bit, and the
    oracle within the bar; visible_lists recognises that list object only (not a copy, not after an in-place
    change), and the drop-in SH_eval takes the geometry-state path for it with the same values and gradients."""
    import torch
    from oracle import oracle as O
    from sv_raster_amd import _C, renderer, synthetic as S
    sc = S.make_scene("A", seed=3)
    cam = S.make_camera(800, 800, seed=3)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    sh0, shs = t(sc["sh0"]), t(sc["shs"])
    N = centers.shape[0]
    w2c, c2w = t(cam["w2c"]), t(cam["c2w"])
    print(f"sh in-frustum list deg {deg}", flush=True)
    # a narrowed field of view leaves part of the voxels outside the frustum (unlisted rows)
    _, geom = _C.rasterize_preprocess(800, 800, 0.3 * cam["tanfovx"], 0.3 * cam["tanfovy"], cam["cx"], cam["cy"], w2c,
                                      c2w, 0.02, octpath, centers, lengths)
    idx, _ = _C.visible_voxels(geom, N, centers.device, _C.order_table_cache.get(octpath))
    assert 0 < idx.numel() < N
    assert _C.visible_lists.geom_of(idx, N) is geom
    assert _C.visible_lists.geom_of(idx.clone(), N) is None
    assert _C.visible_lists.geom_of(idx, N + 1) is None
    cam_pos = c2w[:3, 3].contiguous()
    rgbs_list = _C.sh_compute(deg, idx.clone(), centers, cam_pos, sh0, shs)
    rgbs_vis = _C.sh_compute(deg, idx, centers, cam_pos, sh0, shs, visible=geom)
    assert torch.equal(rgbs_list, rgbs_vis)
    ref = O.sh_fwd(deg, idx.cpu().numpy(), sc["vox_centers"], cam["c2w"][:3, 3], sc["sh0"], sc["shs"])
    assert np.abs(rgbs_vis.cpu().numpy() - ref).max() <= RTOL * max(1.0, np.abs(ref).max())
    dL = t(np.random.default_rng(deg).normal(0, 1, (N, 3)).astype(np.float32))
    g_list = _C.sh_compute_bw(deg, 16, idx.clone(), centers, cam_pos, rgbs_list, dL)
    g_vis = _C.sh_compute_bw(deg, 16, idx, centers, cam_pos, rgbs_vis, dL, visible=geom)
    for a, b in zip(g_list, g_vis):
        assert torch.equal(a, b)
    # the drop-in: the recognised list takes the geometry-state path, a copy the index path; same results
    outs = []
    for sel in (idx, idx.clone()):
        s0, s1 = sh0.clone().requires_grad_(), shs.clone().requires_grad_()
        rgbs = renderer.SH_eval.apply(deg, sel, centers, cam_pos, None, s0, s1)
        (rgbs * dL).sum().backward()
        outs.append((rgbs.detach(), s0.grad, s1.grad))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    # an in-place change of the list makes it unknown
    idx[0] = idx[0]
    assert _C.visible_lists.geom_of(idx, N) is None
