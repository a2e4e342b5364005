"""ORACLE — test infrastructure, not product code.

numpy/ctypes front end of the CPU restatement in ``svr_oracle.cpp``.  Only the
tests, ``__graft_entry__.smoke()`` and the benchmark's ``cpu_baseline`` leg may
import this module, and only as the checker / the timed CPU baseline.

Every function mirrors one reference entry point (cited per function); inputs
are numpy arrays in the reference's layouts (centres [N,3], lengths [N] or
[N,1], geos [N,8], rgbs [N,3], images planar [3,H,W], matrices row-major 4x4).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsvr_oracle.so")
_lib = None

BLOCK = 16


def build():
    """Compile the oracle with its Makefile (g++ -O3 -fopenmp -ffp-contract=off)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_expf.restype = ctypes.c_float
        _lib.oracle_expf.argtypes = [ctypes.c_float]
        _lib.oracle_num_threads.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_num_threads(n):
    lib().oracle_set_num_threads(ctypes.c_int(int(n)))


def num_threads():
    return int(lib().oracle_num_threads())


def expf(x):
    x = _f32(x).ravel()
    y = np.empty_like(x)
    lib().oracle_expf_array(ctypes.c_int(x.size), _p(x), _p(y))
    return y


def tile_grid(W, H):
    return (W + BLOCK - 1) // BLOCK, (H + BLOCK - 1) // BLOCK


def preprocess(cam, centers, lengths):
    """NC/src/preprocess.cu:24-199 -> (n_dup u32 [N], bbox u32 [N,2], qbits u32 [N])."""
    centers = _f32(centers).reshape(-1, 3)
    lengths = _f32(lengths).reshape(-1)
    N = centers.shape[0]
    n_dup = np.zeros(N, np.uint32)
    bbox = np.zeros((N, 2), np.uint32)
    qbits = np.zeros(N, np.uint32)
    w2c = _f32(cam["w2c"]).ravel()
    c2w = _f32(cam["c2w"]).ravel()
    lib().oracle_preprocess(
        ctypes.c_int(cam["W"]), ctypes.c_int(cam["H"]),
        ctypes.c_float(cam["tanfovx"]), ctypes.c_float(cam["tanfovy"]),
        ctypes.c_float(cam["cx"]), ctypes.c_float(cam["cy"]),
        _p(w2c), _p(c2w), ctypes.c_float(cam.get("near", 0.02)),
        ctypes.c_int(N), _p(centers), _p(lengths), _p(n_dup), _p(bbox), _p(qbits))
    return n_dup, bbox, qbits


def binning(cam, octpaths, n_dup, bbox, qbits):
    """NC/src/forward.cu:523-753 -> (keys_hi u64 [R], keys_lo u64 [R], vox_list u32 [R], ranges u32 [T,2])."""
    octpaths = np.ascontiguousarray(octpaths, dtype=np.int64).reshape(-1)
    N = octpaths.size
    R = int(n_dup.astype(np.int64).sum())
    gx, gy = tile_grid(cam["W"], cam["H"])
    hi = np.zeros(R, np.uint64)
    lo = np.zeros(R, np.uint64)
    vl = np.zeros(R, np.uint32)
    ranges = np.zeros((gx * gy, 2), np.uint32)
    lib().oracle_binning(
        ctypes.c_int(cam["W"]), ctypes.c_int(cam["H"]), ctypes.c_int(N), _p(octpaths),
        _p(np.ascontiguousarray(n_dup, np.uint32)), _p(np.ascontiguousarray(bbox, np.uint32)),
        _p(np.ascontiguousarray(qbits, np.uint32)), ctypes.c_int64(R),
        _p(hi), _p(lo), _p(vl), _p(ranges))
    return hi, lo, vl, ranges


def render_fwd(cam, ranges, vox_list, bbox, centers, lengths, geos, rgbs, bg=0.0, n_samp=1, dtype=np.float32):
    """NC/src/forward.cu:43-394 -> (color [3,H,W], T [H,W], n_contrib u32 [H,W], tile_last u32 [T])."""
    W, H = cam["W"], cam["H"]
    gx, gy = tile_grid(W, H)
    cv = _f64 if dtype == np.float64 else _f32
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    centers, lengths, geos, rgbs = cv(centers), cv(lengths).reshape(-1), cv(geos), cv(rgbs)
    c2w = cv(cam["c2w"]).ravel()
    color = np.zeros((3, H, W), dtype)
    T = np.zeros((H, W), dtype)
    nc = np.zeros((H, W), np.uint32)
    tl = np.zeros(gx * gy, np.uint32)
    fn = lib().oracle_render_fwd_f64 if dtype == np.float64 else lib().oracle_render_fwd_f32
    fn(ctypes.c_int(W), ctypes.c_int(H), ct(cam["tanfovx"]), ct(cam["tanfovy"]), ct(cam["cx"]), ct(cam["cy"]),
       _p(c2w), ct(bg), ctypes.c_int(n_samp),
       _p(np.ascontiguousarray(ranges, np.uint32)), _p(np.ascontiguousarray(vox_list, np.uint32)),
       _p(np.ascontiguousarray(bbox, np.uint32)),
       _p(centers), _p(lengths), _p(geos), _p(rgbs), _p(color), _p(T), _p(nc), _p(tl))
    return color, T, nc, tl


def render_bwd(cam, ranges, vox_list, bbox, centers, lengths, geos, rgbs, out_T, n_contrib, tile_last,
               dL_dcolor, dL_dT, bg=0.0, n_samp=1, dtype=np.float32):
    """NC/src/backward.cu:37-772 -> (dL_dgeos [N,8], dL_drgbs [N,3], subdiv_p_bw [N,1]) in float64."""
    W, H = cam["W"], cam["H"]
    cv = _f64 if dtype == np.float64 else _f32
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    centers, lengths, geos, rgbs = cv(centers), cv(lengths).reshape(-1), cv(geos), cv(rgbs)
    N = centers.shape[0]
    c2w = cv(cam["c2w"]).ravel()
    acc = np.zeros((N, 12), np.float64)
    fn = lib().oracle_render_bwd_f64 if dtype == np.float64 else lib().oracle_render_bwd_f32
    fn(ctypes.c_int(W), ctypes.c_int(H), ct(cam["tanfovx"]), ct(cam["tanfovy"]), ct(cam["cx"]), ct(cam["cy"]),
       _p(c2w), ct(bg), ctypes.c_int(n_samp),
       _p(np.ascontiguousarray(ranges, np.uint32)), _p(np.ascontiguousarray(vox_list, np.uint32)),
       _p(np.ascontiguousarray(bbox, np.uint32)),
       _p(centers), _p(lengths), _p(geos), _p(rgbs), _p(cv(out_T)),
       _p(np.ascontiguousarray(n_contrib, np.uint32)), _p(np.ascontiguousarray(tile_last, np.uint32)),
       _p(cv(dL_dcolor)), _p(cv(dL_dT)), ctypes.c_int(N), _p(acc))
    return acc[:, :8].copy(), acc[:, 8:11].copy(), acc[:, 11:12].copy()


def render_fwd_ext(cam, ranges, vox_list, bbox, centers, lengths, geos, rgbs, bg=0.0, n_samp=1, need_depth=False,
                   need_dist=False, need_normal=False, track_max_w=False, dist_loss=False, dtype=np.float32):
    """NC/src/forward.cu:43-394 with the optional outputs -> dict(color [3,H,W], T [H,W], n_contrib, tile_last,
    depth [3,H,W] (mean, distortion sum, median), normal [3,H,W], max_w [N], dist_loss [H,W] (the per-pixel
    distortion loss value, oracle only: sum_ij w_i w_j |m_i - m_j| + 1/3 sum_i w_i^2 span_i))."""
    W, H = cam["W"], cam["H"]
    gx, gy = tile_grid(W, H)
    cv = _f64 if dtype == np.float64 else _f32
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    centers, lengths, geos, rgbs = cv(centers), cv(lengths).reshape(-1), cv(geos), cv(rgbs)
    N = centers.shape[0]
    c2w = cv(cam["c2w"]).ravel()
    color = np.zeros((3, H, W), dtype)
    T = np.zeros((H, W), dtype)
    nc = np.zeros((H, W), np.uint32)
    tl = np.zeros(gx * gy, np.uint32)
    depth = np.zeros((3, H, W), dtype)
    normal = np.zeros((3, H, W), dtype)
    max_w = np.zeros(N, dtype)
    dl = np.zeros((H, W), dtype)
    fn = lib().oracle_render_fwd_ext_f64 if dtype == np.float64 else lib().oracle_render_fwd_ext_f32
    fn(ctypes.c_int(W), ctypes.c_int(H), ct(cam["tanfovx"]), ct(cam["tanfovy"]), ct(cam["cx"]), ct(cam["cy"]),
       _p(c2w), ct(bg), ctypes.c_int(n_samp),
       _p(np.ascontiguousarray(ranges, np.uint32)), _p(np.ascontiguousarray(vox_list, np.uint32)),
       _p(np.ascontiguousarray(bbox, np.uint32)), _p(centers), _p(lengths), _p(geos), _p(rgbs),
       ctypes.c_int(int(need_depth)), ctypes.c_int(int(need_dist)), ctypes.c_int(int(need_normal)),
       ctypes.c_int(int(track_max_w)), _p(color), _p(T), _p(nc), _p(tl), _p(depth), _p(normal), _p(max_w),
       _p(dl) if dist_loss else ctypes.c_void_p(0))
    return dict(color=color, T=T, n_contrib=nc, tile_last=tl, depth=depth, normal=normal, max_w=max_w,
                dist_loss=dl)


def render_bwd_ext(cam, ranges, vox_list, bbox, centers, lengths, geos, rgbs, out_T, n_contrib, tile_last,
                   dL_dcolor, dL_dT, bg=0.0, n_samp=1, need_depth=False, need_normal=False, dL_ddepth=None,
                   dL_dnormal=None, lambda_R=0.0, gt=None, lambda_asc=0.0, lambda_dist=0.0, out_depth=None,
                   dtype=np.float32):
    """NC/src/backward.cu:37-772 with the depth / normal / distortion gradients and the R_concen / ascending
    regularizers -> (dL_dgeos [N,8], dL_drgbs [N,3], subdiv_p_bw [N,1]) in float64."""
    W, H = cam["W"], cam["H"]
    cv = _f64 if dtype == np.float64 else _f32
    ct = ctypes.c_double if dtype == np.float64 else ctypes.c_float
    centers, lengths, geos, rgbs = cv(centers), cv(lengths).reshape(-1), cv(geos), cv(rgbs)
    N = centers.shape[0]
    c2w = cv(cam["c2w"]).ravel()
    acc = np.zeros((N, 12), np.float64)
    z3 = np.zeros((3, H, W), dtype)
    opt = lambda a: cv(a) if a is not None else z3  # noqa: E731
    fn = lib().oracle_render_bwd_ext_f64 if dtype == np.float64 else lib().oracle_render_bwd_ext_f32
    fn(ctypes.c_int(W), ctypes.c_int(H), ct(cam["tanfovx"]), ct(cam["tanfovy"]), ct(cam["cx"]), ct(cam["cy"]),
       _p(c2w), ct(bg), ctypes.c_int(n_samp),
       _p(np.ascontiguousarray(ranges, np.uint32)), _p(np.ascontiguousarray(vox_list, np.uint32)),
       _p(np.ascontiguousarray(bbox, np.uint32)), _p(centers), _p(lengths), _p(geos), _p(rgbs), _p(cv(out_T)),
       _p(np.ascontiguousarray(n_contrib, np.uint32)), _p(np.ascontiguousarray(tile_last, np.uint32)),
       _p(cv(dL_dcolor)), _p(cv(dL_dT)), ctypes.c_int(int(need_depth)), ctypes.c_int(int(need_normal)),
       _p(opt(dL_ddepth)), _p(opt(dL_dnormal)), ct(lambda_R), _p(opt(gt)), ct(lambda_asc), ct(lambda_dist),
       _p(opt(out_depth)), ctypes.c_int(N), _p(acc))
    return acc[:, :8].copy(), acc[:, 8:11].copy(), acc[:, 11:12].copy()


def sh_fwd(deg, idx, centers, cam_pos, sh0, shs):
    """NC/src/sh_compute.cu:180-270 -> rgbs [N,3] (zero outside idx)."""
    centers, sh0, shs = _f32(centers), _f32(sh0), _f32(shs)
    N = centers.shape[0]
    M = 1 + shs.shape[1]
    idx = np.ascontiguousarray(idx if idx is not None else np.zeros(0), np.int64)
    cam = _f32(cam_pos).ravel()
    out = np.zeros((N, 3), np.float32)
    lib().oracle_sh_fwd(ctypes.c_int(deg), ctypes.c_int(idx.size), _p(idx), ctypes.c_int(N), _p(centers),
                        _p(cam), _p(sh0), _p(shs), ctypes.c_int(M), _p(out))
    return out


def sh_bwd(deg, M, idx, centers, cam_pos, dL_drgbs):
    """NC/src/sh_compute.cu:212-299 -> (dL_dsh0 [N,3], dL_dshs [N,M-1,3])."""
    centers = _f32(centers)
    N = centers.shape[0]
    idx = np.ascontiguousarray(idx if idx is not None else np.zeros(0), np.int64)
    cam = _f32(cam_pos).ravel()
    g0 = np.zeros((N, 3), np.float32)
    gs = np.zeros((N, M - 1, 3), np.float32)
    lib().oracle_sh_bwd(ctypes.c_int(deg), ctypes.c_int(idx.size), _p(idx), ctypes.c_int(N), _p(centers),
                        _p(cam), _p(_f32(dL_drgbs)), ctypes.c_int(M), _p(g0), _p(gs))
    return g0, gs


def adam_step(param, grad, exp_avg, exp_avg_sq, step, lr, beta1, beta2, eps, biased=False, sparse=False):
    """NC/src/adam_step.cu:21-162; updates param/exp_avg/exp_avg_sq (float32 arrays) in place."""
    for a in (param, grad, exp_avg, exp_avg_sq):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    lib().oracle_adam_step(ctypes.c_int(int(biased)), ctypes.c_int(int(sparse)), ctypes.c_int64(param.size),
                           _p(param), _p(grad), _p(exp_avg), _p(exp_avg_sq), ctypes.c_double(step),
                           ctypes.c_double(lr), ctypes.c_double(beta1), ctypes.c_double(beta2), ctypes.c_float(eps))


def voxel_order_rank(octpaths):
    """NC/src/utils.cu:106-127 -> int64 [8, N]."""
    octpaths = np.ascontiguousarray(octpaths, np.int64).reshape(-1)
    out = np.zeros((8, octpaths.size), np.int64)
    lib().oracle_voxel_order_rank(ctypes.c_int(octpaths.size), _p(octpaths), _p(out))
    return out


def full_forward(cam, octpaths, centers, lengths, geos, rgbs, bg=0.0, n_samp=1):
    """Preprocess + binning + compositor, as one dict (the reference's forward path)."""
    n_dup, bbox, qbits = preprocess(cam, centers, lengths)
    hi, lo, vl, ranges = binning(cam, octpaths, n_dup, bbox, qbits)
    color, T, nc, tl = render_fwd(cam, ranges, vl, bbox, centers, lengths, geos, rgbs, bg=bg, n_samp=n_samp)
    return dict(n_dup=n_dup, bbox=bbox, qbits=qbits, keys_hi=hi, keys_lo=lo, vox_list=vl, ranges=ranges,
                color=color, T=T, n_contrib=nc, tile_last=tl)


def gather(vox_key, care_idx, grid, n_dim=1):
    """NC/src/geo_params_gather.cu:22-108 -> float32 [N, 8] (n_dim 1) or [N, 8, n_dim]; rows outside care are 0."""
    vox_key = np.ascontiguousarray(vox_key, np.int64)
    care_idx = np.ascontiguousarray(care_idx, np.int64).reshape(-1)
    grid = _f32(grid)
    N = vox_key.shape[0]
    out = np.zeros((N, 8, n_dim), np.float32)
    lib().oracle_gather(ctypes.c_int64(care_idx.size), ctypes.c_int(n_dim), _p(vox_key), _p(care_idx), _p(grid),
                        _p(out))
    return out.reshape(N, 8) if n_dim == 1 else out


def gather_bw(vox_key, care_idx, n_grid, dL, n_dim=1):
    """NC/src/geo_params_gather.cu:47-132 -> float64 [n_grid, n_dim]."""
    vox_key = np.ascontiguousarray(vox_key, np.int64)
    care_idx = np.ascontiguousarray(care_idx, np.int64).reshape(-1)
    out = np.zeros((n_grid, n_dim), np.float64)
    lib().oracle_gather_bw(ctypes.c_int64(care_idx.size), ctypes.c_int(n_dim), _p(vox_key), _p(care_idx),
                           _p(_f32(dL)), _p(out))
    return out


def tv_bw(grid, vox_key, weight, vox_size_inv, no_tv_s, tv_sparse, grad):
    """NC/src/tv_compute.cu:23-80 -> float64 [M, C]: grad plus the TV gradient."""
    grid = _f32(grid)
    M, C = grid.shape
    vox_key = np.ascontiguousarray(vox_key, np.int64)
    out = np.ascontiguousarray(grad, np.float64).copy()
    vsi = _f32(vox_size_inv).reshape(-1)
    lib().oracle_tv_bw(ctypes.c_int64(vox_key.shape[0]), ctypes.c_int(C), _p(grid), _p(vox_key),
                       ctypes.c_float(weight), _p(vsi), ctypes.c_int(int(no_tv_s)), ctypes.c_int(int(tv_sparse)),
                       _p(out))
    return out


def is_in_cone(tanfovx, tanfovy, near, w2c, pts):
    """NC/src/utils.cu:23-43 -> bool [n]."""
    pts = _f32(pts)
    out = np.zeros(pts.shape[0], np.uint8)
    lib().oracle_is_in_cone(ctypes.c_int64(pts.shape[0]), ctypes.c_float(tanfovx), ctypes.c_float(tanfovy),
                            ctypes.c_float(near), _p(_f32(w2c)), _p(pts), _p(out))
    return out.astype(bool)


def pixel_rays(W, H, cx, cy, tanfovx, tanfovy, c2w, depth=None):
    """NC/src/utils.cu:45-102 -> float32 [3, H, W] (compute_rd, or depth2pts with depth)."""
    out = np.zeros((3, H, W), np.float32)
    d = _f32(depth).reshape(-1) if depth is not None else None
    lib().oracle_pixel_rays(ctypes.c_int(W), ctypes.c_int(H), ctypes.c_float(cx), ctypes.c_float(cy),
                            ctypes.c_float(tanfovx), ctypes.c_float(tanfovy), _p(_f32(c2w)),
                            _p(d) if d is not None else ctypes.c_void_p(0), _p(out))
    return out

