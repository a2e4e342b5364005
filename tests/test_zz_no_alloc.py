"""svr_rasterize_backward_noalloc (the backward with its per-entry rows in the
library's fixed device scratch, for callers that cannot allocate) against
svr_rasterize_backward with a caller-allocated workspace, both through the C
ABI on the same forward state.  Two calls in a row, the second on a smaller
frame (fewer entries), so the second reads scratch the first left behind.
Bar: gradients within 1e-4 relative (absolute floor 1e-3 x max|g|): the
per-entry rows are summed in LDS in a run-dependent order."""
import ctypes

import numpy as np
import pytest

from tests._harness import make_case, rel_err

GRAD_RTOL = 1e-4


def _forward(case):
    import torch
    from sv_raster_amd import _C
    sc, cam = case["scene"], case["cam"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    d = dict(octpath=t(sc["octpath"]), centers=t(sc["vox_centers"]), lengths=t(sc["vox_lengths"]),
             geos=t(sc["geos"]), sh0=t(sc["sh0"]), shs=t(sc["shs"]), w2c=t(cam["w2c"]), c2w=t(cam["c2w"]),
             dcol=t(case["dL_dcolor"]), dT=t(case["dL_dT"]))
    W, H = cam["W"], cam["H"]
    N = d["centers"].shape[0]
    n_dup, geom = _C.rasterize_preprocess(W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], d["w2c"],
                                          d["c2w"], cam["near"], d["octpath"], d["centers"], d["lengths"])
    idx, _ = _C.visible_voxels(geom, N, d["centers"].device)
    rgbs = _C.sh_compute(3, idx, d["centers"], d["c2w"][:3, 3], d["sh0"], d["shs"])
    R, binning, img, color, _, _, T, _ = _C.rasterize_voxels(
        1, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], d["w2c"], d["c2w"], 0.0, False, False, False,
        False, d["octpath"], d["centers"], d["lengths"], d["geos"], rgbs, geom)
    d.update(R=R, geom=geom, binning=binning, img=img, T=T, rgbs=rgbs, N=N)
    return d


def _noalloc(case, d):
    import torch
    from sv_raster_amd import _C, _lib
    L = _lib.load()
    cam = case["cam"]
    N = d["N"]
    dg = torch.empty((N, 8), dtype=torch.float32, device="cuda")
    dr = torch.empty((N, 3), dtype=torch.float32, device="cuda")
    ds = torch.empty((N, 1), dtype=torch.float32, device="cuda")
    P = _C._ptr
    rc = L.svr_rasterize_backward_noalloc(1, cam["W"], cam["H"], cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"],
                                          P(d["c2w"]), 0.0, N, P(d["centers"]), P(d["lengths"]), P(d["geos"]),
                                          P(d["rgbs"]), P(d["geom"]), d["R"], P(d["binning"]), P(d["img"]),
                                          P(d["T"]), P(d["dcol"]), P(d["dT"]), P(dg), P(dr), P(ds),
                                          _C._stream(torch.device("cuda")))
    assert rc == 0, L.svr_last_error().decode()
    return dg, dr, ds


def _with_workspace(case, d):
    from sv_raster_amd import _C
    cam = case["cam"]
    return _C.rasterize_voxels_backward(
        d["R"], 1, cam["W"], cam["H"], cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], d["w2c"], d["c2w"], 0.0,
        d["octpath"], d["centers"], d["lengths"], d["geos"], d["rgbs"], d["geom"], d["binning"], d["img"], d["T"],
        d["dcol"], None, None, d["dT"], 0.0, None, 0.0, 0.0, False, False, None, None)


@pytest.mark.gpu
def test_noalloc_backward_matches_workspace_backward_twice():
    import torch
    cases = [("A", make_case("A", seed=2)), ("C", make_case("C", seed=3))]
    for name, case in cases:
        d = _forward(case)
        print(f"noalloc {name}: R = {d['R']}", flush=True)
        got = _noalloc(case, d)
        want = _with_workspace(case, d)
        torch.cuda.synchronize()
        for g, w, what in zip(got, want, ("dgeos", "drgbs", "dsub")):
            e = rel_err(g.cpu().numpy(), w.cpu().numpy())
            assert e <= GRAD_RTOL, f"{name} {what}: rel err {e:.3g}"
    assert cases[1][1]["cam"]["W"] * cases[1][1]["cam"]["H"] < cases[0][1]["cam"]["W"] * cases[0][1]["cam"]["H"]


@pytest.mark.gpu
def test_noalloc_backward_rejects_more_entries_than_its_scratch():
    from sv_raster_amd import _lib
    L = _lib.load()
    limit = 1 << 22
    rc = L.svr_rasterize_backward_noalloc(1, 64, 64, 0.5, 0.5, 32.0, 32.0, None, 0.0, 10, None, None, None, None,
                                          None, limit + 1, None, None, None, None, None, None, None, None, None)
    assert rc == 4  # SVR_ERR_CAPACITY
    assert "exceed" in L.svr_last_error().decode()
