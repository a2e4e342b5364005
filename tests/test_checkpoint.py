"""Checkpoint reader (the reference's SVInOut format): CPU checks of the file
handling (quantized blocks, layout / backend checks, round trip) and a GPU
check that a loaded model renders what the oracle renders from the same
grid-point densities, with voxel geometry and corner links equal to a numpy
restatement."""
import numpy as np
import pytest
import torch

from sv_raster_amd import checkpoint, synthetic as S


def _state(n=50, m=80, k=15, seed=0):
    g = torch.Generator().manual_seed(seed)
    return {"backend_name": "new_cuda", "color_layout": "voxel", "geo_layout": "trilinear", "active_sh_degree": 3,
            "scene_center": torch.zeros(3), "inside_extent": torch.ones(1), "scene_extent": torch.ones(1),
            "octpath": torch.arange(n, dtype=torch.int64).view(-1, 1),
            "octlevel": torch.full((n, 1), 21, dtype=torch.int8),
            "_geo_grid_pts": torch.randn(m, 1, generator=g), "_sh0": torch.rand(n, 3, generator=g),
            "_shs": torch.randn(n, k, 3, generator=g) * 0.05, "quantized": False}


def _quant(t, g):
    """Any index / codebook pair in the file layout (not the reference's codebook fit)."""
    cb = torch.sort(torch.randn(256, generator=g)).values
    idx = torch.randint(0, 256, t.shape, generator=g, dtype=torch.int64).to(torch.uint8)
    return {"index": idx, "codebook": cb}


def test_quantized_blocks_expand_to_codebook_values(tmp_path):
    sd = _state()
    g = torch.Generator().manual_seed(3)
    q = dict(sd, quantized=True)
    q["_geo_grid_pts"] = _quant(sd["_geo_grid_pts"], g)
    q["_sh0"] = [_quant(v, g) for v in sd["_sh0"].split(1, dim=1)]
    q["_shs"] = [_quant(v, g) for v in sd["_shs"].split(1, dim=1)]
    path = str(tmp_path / "q.pt")
    torch.save(q, path)
    out = checkpoint.read_state(path)
    assert torch.equal(out["_geo_grid_pts"], q["_geo_grid_pts"]["codebook"][q["_geo_grid_pts"]["index"].long()])
    assert out["_sh0"].shape == (50, 3) and out["_shs"].shape == (50, 15, 3)
    for c in range(3):
        blk = q["_sh0"][c]
        assert torch.equal(out["_sh0"][:, c:c + 1], blk["codebook"][blk["index"].long()])
    for c in range(15):
        blk = q["_shs"][c]
        assert torch.equal(out["_shs"][:, c:c + 1], blk["codebook"][blk["index"].long()])


def test_layout_and_backend_checks(tmp_path):
    path = str(tmp_path / "m.pt")
    torch.save(dict(_state(), backend_name="new_cuda_cont"), path)
    with pytest.raises(RuntimeError, match="does not match model backend 'new_cuda'"):
        checkpoint.read_state(path)
    torch.save(dict(_state(), color_layout="mixed"), path)
    with pytest.raises(RuntimeError, match="layout"):
        checkpoint.read_state(path)
    st = _state()
    del st["backend_name"]  # files written before the field existed
    torch.save(st, path)
    assert checkpoint.read_state(path)["active_sh_degree"] == 3


def test_save_read_round_trip(tmp_path):
    st = _state(seed=5)
    path = str(tmp_path / "sub" / "iter000100_model.pt")
    checkpoint.save(path, st)
    out = checkpoint.read_state(path)
    for k, v in st.items():
        if torch.is_tensor(v):
            assert torch.equal(out[k], v), k
        else:
            assert out[k] == v, k


class _Cam:
    def __init__(self, cam):
        self.image_width, self.image_height = cam["W"], cam["H"]
        self.tanfovx, self.tanfovy, self.cx, self.cy = cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"]
        self.w2c = torch.from_numpy(cam["w2c"]).cuda()
        self.c2w = torch.from_numpy(cam["c2w"]).cuda()
        self.near = cam["near"]


def _np_link(octpath, octlevel):
    lv = octlevel.reshape(-1).astype(np.int64)
    ijk = S.octpath_to_ijk(octpath, lv).astype(np.int64)
    sh = (21 - lv)[:, None, None]
    corner = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], np.int64)[None]
    pts = (ijk[:, None, :] << sh) + (corner << sh)
    key, inv = np.unique(pts.reshape(-1, 3), axis=0, return_inverse=True)
    return key, inv.reshape(-1, 8)


@pytest.mark.gpu
def test_loaded_model_renders_like_the_oracle(tmp_path):
    from oracle import oracle as O
    tree = S.make_octree(1500, 3, 6, seed=4)
    key, vox_key = _np_link(tree["octpath"], tree["octlevel"])
    rng = np.random.default_rng(4)
    n, m = tree["octpath"].shape[0], key.shape[0]
    st = {"active_sh_degree": 3, "scene_center": torch.from_numpy(np.asarray(tree["scene_center"], np.float32)),
          "inside_extent": torch.tensor([float(tree["scene_extent"])], dtype=torch.float32),
          "scene_extent": torch.tensor([float(tree["scene_extent"])], dtype=torch.float32),
          "octpath": torch.from_numpy(tree["octpath"].reshape(-1, 1).astype(np.int64)),
          "octlevel": torch.from_numpy(tree["octlevel"].reshape(-1, 1).astype(np.int8)),
          "_geo_grid_pts": torch.from_numpy(rng.normal(-1, 1, (m, 1)).astype(np.float32)),
          "_sh0": torch.from_numpy(rng.uniform(0, 1, (n, 3)).astype(np.float32)),
          "_shs": torch.from_numpy(rng.normal(0, 0.05, (n, 15, 3)).astype(np.float32))}
    path = str(tmp_path / "model.pt")
    checkpoint.save(path, st)
    model = checkpoint.load(path)
    print("checkpoint model", n, "voxels", m, "grid points", flush=True)
    centers, sizes = S.decode(tree["octpath"], tree["octlevel"], tree["scene_center"], tree["scene_extent"])
    assert np.array_equal(model.vox_center.cpu().numpy(), centers)
    assert np.array_equal(model.vox_size.cpu().numpy(), sizes)
    assert np.array_equal(model.grid_pts_key.cpu().numpy(), key)
    assert np.array_equal(model.vox_key.cpu().numpy(), vox_key)

    cam = S.make_camera(96, 72, seed=9)
    out = model.render(_Cam(cam), output_depth=True)
    geos = st["_geo_grid_pts"].numpy()[vox_key, 0]
    n_dup, bbox, qbits = O.preprocess(cam, centers, sizes)
    idx = np.nonzero(n_dup > 0)[0]
    rgbs = O.sh_fwd(3, idx, centers, cam["c2w"][:3, 3], st["_sh0"].numpy(), st["_shs"].numpy())
    _, _, vl, ranges = O.binning(cam, tree["octpath"], n_dup, bbox, qbits)
    e = O.render_fwd_ext(cam, ranges, vl, bbox, centers, sizes, geos, rgbs, need_depth=True)
    color = out["color"].detach().cpu().numpy()
    assert np.abs(color - e["color"]).max() <= 1e-5 * max(1.0, np.abs(e["color"]).max())
    assert np.array_equal(out["depth"].detach().cpu().numpy()[0], e["depth"][0])
