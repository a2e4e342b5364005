"""Batched per-voxel statistics over many cameras (compute_training_stat,
mark_max_samp_rate, mark_near) vs the per-camera loops they replace, run
through the drop-in single-view path exactly as the reference writes them
(SRC/sparse_voxel_gears/adaptive.py:190-221, NC renderer.py:408-438), and
max_w vs the oracle.

Bar: max_w, view_cnt and is_near bit-exact; min_samp_interval and
max_samp_rate within 1e-6 relative (the camera-axis distance is a 3-term
float sum that torch may associate differently)."""
import numpy as np
import pytest

from tests.test_gpu_batch import _settings, _views


class _Cam:
    """The attributes of the reference's CameraBase (cameras.py:29-63) that these functions read."""

    def __init__(self, cam, t):
        self.image_width, self.image_height = cam["W"], cam["H"]
        self.tanfovx, self.tanfovy, self.cx, self.cy = cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"]
        self.w2c, self.c2w = t(cam["w2c"]), t(cam["c2w"])

    @property
    def position(self):
        return self.c2w[:3, 3]

    @property
    def lookat(self):
        return self.c2w[:3, 2]

    @property
    def pix_size(self):
        return 2 * self.tanfovx / self.image_width


def _scene(seed):
    import torch
    specs = [((64, 64), None), ((48, 40), None), ((80, 56), None), ((63, 63), (0.0, 0.0, -1.0)), ((96, 32), None)]
    cases = _views("C", specs, seed=seed)
    sc = cases[0]["scene"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    return cases, sc, t


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30))) if a.size else 0.0


@pytest.mark.gpu
def test_training_stat_matches_per_camera_loop_and_oracle():
    import torch
    from oracle import oracle as O
    from sv_raster_amd import batch, renderer
    cases, sc, t = _scene(11)
    octpath, centers, lengths, geos = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"]), t(sc["geos"])
    N = centers.shape[0]
    print("training stat, C x5", flush=True)
    views = batch.ViewBatch([_settings(c["cam"]) for c in cases])
    got = batch.compute_training_stat(views, octpath, centers, lengths, geos)

    # adaptive.py:199-215, one render per camera
    max_w = torch.zeros([N, 1], dtype=torch.float32, device="cuda")
    msi = torch.full([N, 1], 1e30, dtype=torch.float32, device="cuda")
    cnt = torch.zeros([N, 1], dtype=torch.float32, device="cuda")
    zeros_rgb = torch.zeros((N, 3), device="cuda")

    def vox_fn(idx, cam_pos, color_mode):
        return {"geos": geos, "rgbs": zeros_rgb, "subdiv_p": torch.ones((N, 1), device="cuda")}

    oracle_w = np.zeros(N, np.float32)
    for c in cases:
        cam = _Cam(c["cam"], t)
        s = _settings(c["cam"])._replace(track_max_w=True)
        max_w_i = renderer.rasterize_voxels(s, octpath, centers, lengths, vox_fn)[4]
        max_w = torch.maximum(max_w, max_w_i)
        vis_idx = (max_w_i > 0).squeeze().argwhere().squeeze()
        zdist = ((centers[vis_idx] - cam.position) * cam.lookat).sum(-1, keepdims=True)
        samp_interval = zdist * cam.pix_size
        msi[vis_idx] = torch.minimum(msi[vis_idx], samp_interval)
        cnt[vis_idx] += 1
        n_dup, bbox, qbits = O.preprocess(c["cam"], sc["vox_centers"], sc["vox_lengths"])
        _, _, vl, ranges = O.binning(c["cam"], sc["octpath"], n_dup, bbox, qbits)
        e = O.render_fwd_ext(c["cam"], ranges, vl, bbox, sc["vox_centers"], sc["vox_lengths"], sc["geos"],
                             np.zeros((N, 3), np.float32), track_max_w=True)
        oracle_w = np.maximum(oracle_w, e["max_w"])
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in got.items()}
    assert np.array_equal(g["max_w"], max_w.cpu().numpy())
    assert np.array_equal(g["max_w"].reshape(-1), oracle_w)
    assert np.array_equal(g["view_cnt"], cnt.cpu().numpy())
    assert g["view_cnt"].max() >= 2
    assert _rel(g["min_samp_interval"], msi.cpu().numpy()) <= 1e-6


def _loop_marks(cams, octpath, centers, lengths, near_rate, near_near):
    """The per-camera loops of NC renderer.py:408-438 in torch on the preprocess of the drop-in path: the
    maximum sampling rate over the cameras that see a voxel beyond near_rate, and whether some camera sees it
    within near_near + its length."""
    import torch
    from sv_raster_amd import renderer
    n = centers.shape[0]
    rate = torch.zeros(n, dtype=torch.float32, device=centers.device)
    close = torch.zeros(n, dtype=torch.bool, device=centers.device)
    length = lengths.reshape(-1)
    for cam in cams:
        z = ((centers - cam.position) * cam.lookat).sum(-1)
        seen = renderer.mark_n_duplicates(cam.image_width, cam.image_height, cam.tanfovx, cam.tanfovy, cam.cx, cam.cy,
                                          cam.w2c, cam.c2w, near_rate, octpath, centers, lengths) > 0
        vis = seen & (z > near_rate)
        cand = torch.where(vis, length / (z * cam.pix_size), torch.zeros_like(rate))
        rate = torch.maximum(rate, cand)
        seen = renderer.mark_n_duplicates(cam.image_width, cam.image_height, cam.tanfovx, cam.tanfovy, cam.cx, cam.cy,
                                          cam.w2c, cam.c2w, near_near, octpath, centers, lengths) > 0
        close |= seen & (z <= near_near + length)
    return rate, close


@pytest.mark.gpu
@pytest.mark.parametrize("near", [0.02, 0.2, 1.5])
def test_mark_max_samp_rate_and_near_match_per_camera_loop(near):
    import torch
    from sv_raster_amd import batch, renderer
    cases, sc, t = _scene(12)
    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    cams = [_Cam(c["cam"], t) for c in cases]
    print("mark near", near, flush=True)
    rate_b = renderer.mark_max_samp_rate(cams, octpath, centers, lengths, near=near)
    near_b = renderer.mark_near(cams, octpath, centers, lengths, near=near)
    assert torch.equal(rate_b, batch.mark_max_samp_rate(cams, octpath, centers, lengths, near=near))
    rate_r, near_r = _loop_marks(cams, octpath, centers, lengths, near, near)
    torch.cuda.synchronize()
    assert rate_b.shape == rate_r.shape and near_b.dtype == torch.bool and near_b.shape == near_r.shape
    rb, rr = rate_b.cpu().numpy(), rate_r.cpu().numpy()
    assert np.array_equal(rb > 0, rr > 0) and (rr > 0).any()
    assert _rel(rb, rr) <= 1e-6
    assert np.array_equal(near_b.cpu().numpy(), near_r.cpu().numpy())
