"""Runs the HIP path (through sv_raster_amd._C, i.e. the C ABI) and the oracle
on the same seeded scene, returning numpy arrays for comparison."""
import numpy as np

from sv_raster_amd import synthetic as S


def make_case(name, seed=0, n=None, view=0, cfg=None, direction=None):
    scene = S.make_scene(cfg if cfg is not None else name, seed=seed, n=n)
    views = S.config_views(name, seed) if cfg is None else cfg["views"]
    W, H = views[view]
    cam = S.make_camera(W, H, seed=seed + view, direction=direction)
    rng = np.random.default_rng(seed + 99)
    dL_dcolor = (rng.normal(0, 1, (3, H, W)) / (H * W)).astype(np.float32)
    dL_dT = (rng.normal(0, 1, (1, H, W)) / (H * W)).astype(np.float32)
    return dict(scene=scene, cam=cam, dL_dcolor=dL_dcolor, dL_dT=dL_dT)


def run_oracle(case, bg=0.0, n_samp=1, backward=True):
    from oracle import oracle as O
    sc, cam = case["scene"], case["cam"]
    cam_pos = cam["c2w"][:3, 3]
    n_dup, bbox, qbits = O.preprocess(cam, sc["vox_centers"], sc["vox_lengths"])
    idx = np.nonzero(n_dup > 0)[0]
    rgbs = O.sh_fwd(3, idx, sc["vox_centers"], cam_pos, sc["sh0"], sc["shs"])
    hi, lo, vl, ranges = O.binning(cam, sc["octpath"], n_dup, bbox, qbits)
    color, T, nc, tl = O.render_fwd(cam, ranges, vl, bbox, sc["vox_centers"], sc["vox_lengths"], sc["geos"], rgbs,
                                    bg=bg, n_samp=n_samp)
    out = dict(n_dup=n_dup, bbox=bbox, qbits=qbits, idx=idx, rgbs=rgbs, keys_hi=hi, keys_lo=lo, vox_list=vl,
               ranges=ranges, color=color, T=T, n_contrib=nc, tile_last=tl)
    if backward:
        dgeo, drgb, dsub = O.render_bwd(cam, ranges, vl, bbox, sc["vox_centers"], sc["vox_lengths"], sc["geos"], rgbs,
                                        T, nc, tl, case["dL_dcolor"], case["dL_dT"].reshape(T.shape), bg=bg,
                                        n_samp=n_samp)
        dsh0, dshs = O.sh_bwd(3, 16, idx, sc["vox_centers"], cam_pos, drgb.astype(np.float32))
        out.update(dgeos=dgeo, drgbs=drgb, dsub=dsub, dsh0=dsh0, dshs=dshs)
    return out


def run_gpu(case, bg=0.0, n_samp=1, backward=True, device="cuda"):
    import torch
    from sv_raster_amd import _C
    sc, cam = case["scene"], case["cam"]

    def t(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(device)

    octpath, centers, lengths = t(sc["octpath"]), t(sc["vox_centers"]), t(sc["vox_lengths"])
    geos, sh0, shs = t(sc["geos"]), t(sc["sh0"]), t(sc["shs"])
    w2c, c2w = t(cam["w2c"]), t(cam["c2w"])
    W, H = cam["W"], cam["H"]
    N = centers.shape[0]
    n_dup, geom = _C.rasterize_preprocess(W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w,
                                          cam["near"], octpath, centers, lengths)
    idx, R_vis = _C.visible_voxels(geom, N, centers.device)
    cam_pos = c2w[:3, 3]
    rgbs = _C.sh_compute(3, idx, centers, cam_pos, sh0, shs)
    R, binning, img, color, _, _, T, _ = _C.rasterize_voxels(
        n_samp, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, bg, False, False, False, False,
        octpath, centers, lengths, geos, rgbs, geom)
    assert R == R_vis, (R, R_vis)
    ranges, tile_last, n_contrib = _C.unpack_ImageState(W, H, img)
    hi, lo, vals = _C.export_sort_keys(W, H, R, binning, octpath)
    gnd, gbb, gqb = _C.unpack_geom_state(N, geom)
    out = dict(n_dup=n_dup, geom_n_dup=gnd, bbox=gbb, qbits=gqb, idx=idx, rgbs=rgbs, keys_hi=hi, keys_lo=lo,
               vox_list=vals, ranges=ranges.reshape(-1, 2), tile_last=tile_last.reshape(-1), n_contrib=n_contrib,
               color=color, T=T.reshape(H, W), R=R)
    if backward:
        dgeos, drgbs, dsub = _C.rasterize_voxels_backward(
            R, n_samp, W, H, cam["tanfovx"], cam["tanfovy"], cam["cx"], cam["cy"], w2c, c2w, bg, octpath, centers,
            lengths, geos, rgbs, geom, binning, img, T, t(case["dL_dcolor"]), None, None, t(case["dL_dT"]), 0.0, None,
            0.0, 0.0, False, False, None, None)
        dsh0, dshs = _C.sh_compute_bw(3, 16, idx, centers, cam_pos, rgbs, drgbs)
        out.update(dgeos=dgeos, drgbs=drgbs, dsub=dsub, dsh0=dsh0, dshs=dshs)
    torch.cuda.synchronize()
    res = {}
    for k, v in out.items():
        if torch.is_tensor(v):
            v = v.cpu().numpy()
        res[k] = v
    u32 = lambda a: a.astype(np.int64).astype(np.uint32) if a.dtype != np.uint32 else a  # noqa: E731
    for k in ("geom_n_dup", "bbox", "qbits", "vox_list", "ranges", "tile_last", "n_contrib"):
        res[k] = res[k].view(np.uint32) if res[k].dtype == np.int32 else u32(res[k])
    res["keys_hi"] = res["keys_hi"].view(np.uint64)
    res["keys_lo"] = res["keys_lo"].view(np.uint64)
    return res


def rel_err(a, b, floor_frac=1e-3):
    """max |a-b| / (|b| + floor), floor = floor_frac * max|b| (elementwise relative error
    with a small absolute floor for entries near zero)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.abs(b).max() if b.size else 0.0
    if scale == 0.0:
        return float(np.abs(a).max()) if a.size else 0.0
    diff = np.abs(a - b)
    den = np.abs(b) + floor_frac * scale
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(diff == 0, 0.0, diff / den)
    return float(r.max())
