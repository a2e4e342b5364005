"""HIP gather / total-variation / camera utilities vs the oracle (through the
C ABI): gathers and camera utilities bit-exact (same float operations),
atomically accumulated gradients within the 1e-4 gradient bar (floor 1e-3 x max) of the oracle's float64
sums."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._harness import rel_err
from tests.test_oracle_grid import _grid_case


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
def test_gather_geo_and_feat_match_oracle():
    import torch
    from sv_raster_amd import _C
    rng, vox_key, grid, care = _grid_case(10, N=5000, M=9000)
    M = grid.shape[0]
    out = _C.gather_triinterp_geo_params(_t(vox_key), _t(care), _t(grid[:, 0]))
    ref = O.gather(vox_key, care, grid[:, 0])
    assert np.array_equal(out.cpu().numpy()[care], ref[care])
    dL = rng.normal(0, 1, (vox_key.shape[0], 8)).astype(np.float32)
    g = _C.gather_triinterp_geo_params_bw(_t(vox_key), _t(care), M, _t(dL)).cpu().numpy()
    assert g.shape == (M, 1)
    assert rel_err(g, O.gather_bw(vox_key, care, M, dL)) <= 1e-4
    g3 = rng.normal(0, 1, (M, 3)).astype(np.float32)
    out3 = _C.gather_triinterp_feat_params(_t(vox_key), _t(care), _t(g3))
    assert np.array_equal(out3.cpu().numpy()[care], O.gather(vox_key, care, g3, 3)[care])
    dL3 = rng.normal(0, 1, (vox_key.shape[0], 8, 3)).astype(np.float32)
    gg3 = _C.gather_triinterp_feat_params_bw(_t(vox_key), _t(care), M, _t(dL3)).cpu().numpy()
    assert rel_err(gg3, O.gather_bw(vox_key, care, M, dL3, 3)) <= 1e-4
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_gather_autograd_through_renderer():
    import torch
    from sv_raster_amd import renderer
    rng, vox_key, grid, care = _grid_case(11, N=2000, M=4000)
    pts = _t(grid[:, 0]).requires_grad_()
    geo = renderer.GatherGeoParams.apply(_t(vox_key), _t(care), pts)
    w = _t(rng.normal(0, 1, geo.shape).astype(np.float32))
    (geo[_t(care)] * w[_t(care)]).sum().backward()
    ref = O.gather_bw(vox_key, care, grid.shape[0], w.cpu().numpy())[:, 0]
    assert rel_err(pts.grad.cpu().numpy(), ref) <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("no_tv_s,tv_sparse", [(True, False), (False, False), (True, True), (False, True)])
def test_total_variation_matches_oracle(no_tv_s, tv_sparse):
    from sv_raster_amd import grid_loss_bw
    rng, vox_key, grid, _ = _grid_case(12, N=6000, M=10000, C=2)
    N, M = vox_key.shape[0], grid.shape[0]
    vsi = rng.uniform(1, 60, N).astype(np.float32)
    incoming = np.zeros((M, 2), np.float32)
    hot = rng.choice(M, M // 8, replace=False)
    incoming[hot] = rng.normal(0, 1, (hot.size, 2)).astype(np.float32)
    grad = _t(incoming)
    grid_loss_bw.total_variation(_t(grid), _t(vox_key), 0.3, _t(vsi), no_tv_s, tv_sparse, grad)
    ref = O.tv_bw(grid, vox_key, 0.3, vsi, no_tv_s, tv_sparse, incoming)
    assert rel_err(grad.cpu().numpy(), ref) <= 1e-4


@pytest.mark.gpu
def test_camera_utilities_match_oracle():
    import torch
    from sv_raster_amd import _C, utils
    from sv_raster_amd import synthetic as S
    rng = np.random.default_rng(13)
    cam = S.make_camera(37, 29, seed=3)
    pts = rng.uniform(-3, 3, (20000, 3)).astype(np.float32)
    got = _C.is_in_cone(cam["tanfovx"], cam["tanfovy"], cam["near"], _t(cam["w2c"]), _t(pts)).cpu().numpy()
    assert np.array_equal(got, O.is_in_cone(cam["tanfovx"], cam["tanfovy"], cam["near"], cam["w2c"], pts))
    rd = utils.compute_rd(37, 29, cam["cx"], cam["cy"], cam["tanfovx"], cam["tanfovy"], _t(cam["c2w"]))
    assert np.array_equal(rd.cpu().numpy(), O.pixel_rays(37, 29, cam["cx"], cam["cy"], cam["tanfovx"],
                                                          cam["tanfovy"], cam["c2w"]))
    depth = rng.uniform(0.1, 4, (29, 37)).astype(np.float32)
    p = utils.depth2pts(37, 29, cam["cx"], cam["cy"], cam["tanfovx"], cam["tanfovy"], _t(cam["c2w"]), _t(depth))
    assert np.array_equal(p.cpu().numpy(), O.pixel_rays(37, 29, cam["cx"], cam["cy"], cam["tanfovx"],
                                                         cam["tanfovy"], cam["c2w"], depth))
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_filter_geom_state_selects_voxels():
    import torch
    from sv_raster_amd import _C
    from tests._harness import make_case
    case = make_case("C", seed=1)
    sc, cam = case["scene"], case["cam"]
    N = sc["octpath"].shape[0]
    n_dup, geom = _C.rasterize_preprocess(cam["W"], cam["H"], cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"],
                                          _t(cam["w2c"]), _t(cam["c2w"]), cam["near"], _t(sc["octpath"]),
                                          _t(sc["vox_centers"]), _t(sc["vox_lengths"]))
    keep = np.sort(np.random.default_rng(0).choice(N, N // 3, replace=False)).astype(np.int64)
    new = _C.filter_geomState(N, _t(keep), geom)
    a = [x.cpu().numpy() for x in _C.unpack_geom_state(N, geom)]
    b = [x.cpu().numpy() for x in _C.unpack_geom_state(keep.size, new)]
    for x, y in zip(a, b):
        assert np.array_equal(x[keep], y)
    torch.cuda.synchronize()
