"""The oracle's grid-point gather, total-variation gradient and camera
utilities pinned by library routines and closed forms (no GPU)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O


def _grid_case(seed, N=300, M=500, C=1):
    rng = np.random.default_rng(seed)
    # the 8 corners of a voxel are distinct grid points (as in an octree's corner table)
    vox_key = np.stack([rng.choice(M, 8, replace=False) for _ in range(N)]).astype(np.int64)
    grid = rng.normal(0, 1, (M, C)).astype(np.float32)
    care = np.sort(rng.choice(N, N // 2, replace=False)).astype(np.int64)
    return rng, vox_key, grid, care


def test_gather_is_fancy_indexing():
    rng, vox_key, grid, care = _grid_case(0)
    out = O.gather(vox_key, care, grid[:, 0])
    assert np.array_equal(out[care], grid[:, 0][vox_key[care]])
    rest = np.setdiff1d(np.arange(vox_key.shape[0]), care)
    assert not out[rest].any()
    g3 = rng.normal(0, 1, (grid.shape[0], 3)).astype(np.float32)
    out3 = O.gather(vox_key, care, g3, n_dim=3)
    assert np.array_equal(out3[care], g3[vox_key[care]])


def test_gather_bw_is_scatter_add():
    rng, vox_key, grid, care = _grid_case(1)
    M = grid.shape[0]
    dL = rng.normal(0, 1, (vox_key.shape[0], 8)).astype(np.float32)
    ref = np.zeros(M)
    np.add.at(ref, vox_key[care].reshape(-1), dL[care].reshape(-1).astype(np.float64))
    assert np.allclose(O.gather_bw(vox_key, care, M, dL)[:, 0], ref, rtol=0, atol=1e-12)
    dL3 = rng.normal(0, 1, (vox_key.shape[0], 8, 3)).astype(np.float32)
    ref3 = np.zeros((M, 3))
    np.add.at(ref3, vox_key[care].reshape(-1), dL3[care].reshape(-1, 3).astype(np.float64))
    assert np.allclose(O.gather_bw(vox_key, care, M, dL3, n_dim=3), ref3, rtol=0, atol=1e-12)


def _edge_l1_grad(grid, vox_key, w_per_vox):
    """d/dgrid of sum_n w_n * sum over the 12 cube edges (a, b) of |v_a - v_b| (torch autograd, float64)."""
    g = torch.tensor(grid, dtype=torch.float64, requires_grad=True)
    key = torch.tensor(vox_key)
    v = g[key]  # [N, 8, C]
    loss = 0.0
    for a in range(8):
        for bit in (1, 2, 4):
            b = a ^ bit
            if b > a:
                loss = loss + (torch.as_tensor(w_per_vox, dtype=torch.float64)[:, None] * (v[:, a] - v[:, b]).abs()).sum()
    loss.backward()
    return g.grad.numpy()


@pytest.mark.parametrize("no_tv_s", [True, False])
def test_tv_is_gradient_of_edge_l1(no_tv_s):
    rng, vox_key, grid, _ = _grid_case(2, C=2)
    N = vox_key.shape[0]
    vsi = rng.uniform(1, 50, N).astype(np.float32)
    weight = 0.37
    w = np.full(N, np.float32(weight)) if no_tv_s else (np.float32(weight) * (np.float32(0.01) * vsi))
    ref = _edge_l1_grad(grid, vox_key, w.astype(np.float64))
    got = O.tv_bw(grid, vox_key, weight, vsi, no_tv_s, False, np.zeros_like(grid, np.float64))
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-6)  # terms w * sign are rounded to float


def test_tv_sparse_skips_items_without_incoming_gradient():
    rng, vox_key, grid, _ = _grid_case(3, C=1)
    M, N = grid.shape[0], vox_key.shape[0]
    incoming = np.zeros((M, 1))
    hot = rng.choice(M, M // 10, replace=False)
    incoming[hot, 0] = rng.normal(0, 1, hot.size)
    vsi = np.ones(N, np.float32)
    take = (incoming[vox_key, 0] != 0).any(1)
    ref = incoming + _edge_l1_grad(grid, vox_key, np.where(take, 0.5, 0.0))
    got = O.tv_bw(grid, vox_key, 0.5, vsi, True, True, incoming)
    assert 0 < take.sum() < N
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-6)  # terms w * sign are rounded to float


def test_camera_utilities_closed_form():
    rng = np.random.default_rng(4)
    w2c = np.eye(4, dtype=np.float32)
    w2c[:3, 3] = [0.1, -0.2, 0.3]
    pts = rng.uniform(-2, 2, (5000, 3)).astype(np.float32)
    got = O.is_in_cone(0.5, 0.4, 0.05, w2c, pts)
    cam = pts.astype(np.float64) + w2c[:3, 3]
    ref = (cam[:, 2] > 0.05) & (np.abs(cam[:, 0]) <= 0.5 * cam[:, 2]) & (np.abs(cam[:, 1]) <= 0.4 * cam[:, 2])
    margin = np.minimum(np.abs(np.abs(cam[:, 0]) - 0.5 * cam[:, 2]), np.abs(np.abs(cam[:, 1]) - 0.4 * cam[:, 2]))
    ok = margin > 1e-5
    assert np.array_equal(got[ok], ref[ok])
    c2w = np.eye(4, dtype=np.float32)
    c2w[:3, :3] = np.linalg.qr(rng.normal(size=(3, 3)))[0].astype(np.float32)
    c2w[:3, 3] = [1.0, 2.0, 3.0]
    W, H, cx, cy, tx, ty = 7, 5, 3.3, 2.1, 0.6, 0.45
    rd = O.pixel_rays(W, H, cx, cy, tx, ty, c2w)
    ys, xs = np.mgrid[0:H, 0:W]
    d = np.stack([(xs + 0.5 - cx) * 2 * tx / W, (ys + 0.5 - cy) * 2 * ty / H, np.ones_like(xs, float)])
    ref_rd = np.einsum("ij,jhw->ihw", c2w[:3, :3].astype(np.float64), d)
    assert np.allclose(rd, ref_rd, rtol=1e-5, atol=1e-6)
    depth = rng.uniform(0.5, 3, (H, W)).astype(np.float32)
    pts3 = O.pixel_rays(W, H, cx, cy, tx, ty, c2w, depth)
    assert np.allclose(pts3, c2w[:3, 3, None, None] + depth[None] * ref_rd, rtol=1e-5, atol=1e-5)
