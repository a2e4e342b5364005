"""Tensor-level adapter over the C ABI, with the function names, argument order
and return tuples of the reference extension module `new_svraster_cuda._C`
(backends/new_cuda/binding.cpp:24-57), so that renderer.py / sparse_adam.py /
utils.py read like their reference counterparts.

torch is used only for device memory and the current stream; every compute
step runs in libsvraster_hip.so.
"""
import ctypes
import weakref

import numpy as np

import torch

from . import _lib

_L = _lib.load()

_consts = _lib.SvrConstants()
_lib.check(_L.svr_get_constants(ctypes.byref(_consts)), "svr_get_constants")
MAX_NUM_LEVELS = int(_consts.max_num_levels)
STEP_SZ_SCALE = float(_consts.step_sz_scale)
MAX_N_SAMP = int(_consts.max_n_samp)
MAX_SORT_KEY_BITS = int(_consts.max_sort_key_bits)
MAX_PACKED_IMAGE_DIM = int(_consts.max_packed_image_dim)
MAX_RENDER_TILES = int(_consts.max_render_tiles)
TILE = int(_consts.tile_x)
SVR_FWD_PAIRS_READY = 1  # flags of svr_rasterize_binning / svr_rasterize_forward(_ext), include/svraster_hip.h
SVR_FWD_BINNED = 2


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream(device):
    """The current torch stream of `device` as a hipStream_t."""
    if not isinstance(device, torch.device):
        device = torch.device(device)
    if _raw_stream is not None and device.type == "cuda":
        idx = device.index if device.index is not None else torch.cuda.current_device()
        return ctypes.c_void_p(_raw_stream(idx))
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def _f32(t):
    return t.contiguous() if t.dtype == torch.float32 else t.float().contiguous()


def _f32_a16(t):
    t = _f32(t)
    return t if t.data_ptr() % 16 == 0 else t.clone()


def _require_cuda(*tensors):
    for t in tensors:
        if t is not None and torch.is_tensor(t) and t.numel() > 0 and not t.is_cuda:
            raise RuntimeError("sv_raster_amd needs its tensors on a HIP device (found a CPU tensor)")


def _debug(debug, device, what):
    if debug:
        _lib.check(_L.svr_debug_sync(_stream(device)), what)


def tile_grid(W, H):
    return (W + TILE - 1) // TILE, (H + TILE - 1) // TILE


# ---------------------------------------------------------------------------
# per-octree rank-order table cache

class _OrderTableCache:
    """Caches the rank-order table of the last few octrees, keyed by the
    octree_paths tensor (identity + in-place version counter + storage)."""

    def __init__(self, capacity=4):
        self.capacity = capacity
        self.entries = []

    def clear(self):
        self.entries = []

    def get(self, octree_paths):
        for ref, ver, ptr, n, table in self.entries:
            if ref() is octree_paths and ver == octree_paths._version and ptr == octree_paths.data_ptr() \
                    and n == octree_paths.numel():
                return table
        table = build_order_table(octree_paths)
        self.entries.insert(0, (weakref.ref(octree_paths), octree_paths._version, octree_paths.data_ptr(),
                                octree_paths.numel(), table))
        del self.entries[self.capacity:]
        return table


order_table_cache = _OrderTableCache()


def build_order_table(octree_paths):
    paths = octree_paths.reshape(-1).contiguous()
    if paths.dtype != torch.int64:
        raise RuntimeError("octree_paths must be int64")
    N = paths.numel()
    dev = paths.device
    table = torch.empty(int(_L.svr_order_table_size(N)), dtype=torch.uint8, device=dev)
    if N == 0:
        return table
    scratch = torch.empty(int(_L.svr_order_table_scratch_size(N)), dtype=torch.uint8, device=dev)
    _lib.check(_L.svr_build_order_table(N, _ptr(paths), _ptr(table), _ptr(scratch), _stream(dev)),
               "build_order_table")
    del scratch
    return table


# ---------------------------------------------------------------------------
# binding.cpp:25 -> preprocess.cu:136-199

def rasterize_preprocess(image_width, image_height, tan_fovx, tan_fovy, cx, cy, w2c_matrix, c2w_matrix, near,
                         octree_paths, vox_centers, vox_lengths, debug=False):
    if vox_centers.ndimension() != 2 or vox_centers.size(1) != 3:
        raise RuntimeError("vox_centers must have dimensions (num_points, 3)")
    _require_cuda(w2c_matrix, c2w_matrix, vox_centers, vox_lengths)
    dev = vox_centers.device
    N = vox_centers.size(0)
    out_n_dup = torch.empty(N, dtype=torch.int32, device=dev)  # the preprocess writes every voxel's count
    geom = torch.empty(int(_L.svr_geom_state_size(N)) if N > 0 else 0, dtype=torch.uint8, device=dev)
    if N == 0:
        return out_n_dup, geom
    centers, lengths = _f32(vox_centers), _f32(vox_lengths)
    w2c, c2w = _f32(w2c_matrix), _f32(c2w_matrix)
    _lib.check(_L.svr_rasterize_preprocess(int(image_width), int(image_height), float(tan_fovx), float(tan_fovy),
                                           float(cx), float(cy), _ptr(w2c), _ptr(c2w), float(near), N,
                                           _ptr(centers), _ptr(lengths), _ptr(out_n_dup), _ptr(geom),
                                           _stream(dev)), "rasterize_preprocess")
    _debug(debug, dev, "rasterize_preprocess")
    return out_n_dup, geom


class _VisibleLists:
    """The in-frustum index lists visible_voxels returned lately, each with the geometry state it came from: for
    such a list the SH passes take the listed rows from the state's entry counts (n_dup > 0 lists exactly those
    voxels) instead of zero-filling and marking them first.  A list is recognised by identity, in-place version,
    storage and length; the entries hold weak references, so they keep nothing alive."""

    def __init__(self, capacity=4):
        self.capacity = capacity
        self.entries = []

    def add(self, idx, geom, n_voxels):
        self.entries.insert(0, (weakref.ref(idx), idx._version, idx.data_ptr(), idx.numel(), weakref.ref(geom),
                                int(n_voxels)))
        del self.entries[self.capacity:]

    def geom_of(self, idx, n_voxels):
        """The geometry state whose in-frustum list idx is (for n_voxels voxels), or None."""
        if idx is None or not torch.is_tensor(idx):
            return None
        for ref, ver, ptr, n, gref, nv in self.entries:
            if ref() is idx and idx._version == ver and idx.data_ptr() == ptr and idx.numel() == n \
                    and nv == int(n_voxels):
                return gref()
        return None


visible_lists = _VisibleLists()


def visible_voxels(geomBuffer, n_voxels, device, order_table=None):
    """In-frustum voxel ids (int64, ascending; what renderer.py:81 gets from torch.where(n_duplicates > 0)[0])
    and the number of rendered entries R, from the geometry state of rasterize_preprocess, with one host
    synchronisation for both.  With the octree's order table the binning work that does not need R is queued
    as well; the forward of this geometry state is then called with pairs_ready=True."""
    N = int(n_voxels)
    if N == 0:
        return torch.empty(0, dtype=torch.int64, device=device), 0
    idx = torch.empty(N, dtype=torch.int64, device=device)
    R, nv = ctypes.c_int64(0), ctypes.c_int64(0)
    _lib.check(_L.svr_visible_voxels(N, _ptr(geomBuffer), _ptr(order_table), _ptr(idx), _stream(device),
                                     ctypes.byref(R), ctypes.byref(nv)), "visible_voxels")
    idx = idx[:int(nv.value)]
    visible_lists.add(idx, geomBuffer, N)
    return idx, int(R.value)


def rasterize_binning(image_width, image_height, octree_paths, geomBuffer, num_rendered, pairs_ready=False):
    """The binning half of the forward (key duplication, sort, tile ranges) for R = num_rendered entries, queued
    before the per-voxel colours exist; returns the (binningBuffer, imageBuffer) that rasterize_voxels then
    composites with binned=(binningBuffer, imageBuffer)."""
    dev = geomBuffer.device
    N = octree_paths.numel()
    W, H, R = int(image_width), int(image_height), int(num_rendered)
    img = torch.empty(int(_L.svr_image_state_size(W, H, R)), dtype=torch.uint8, device=dev)
    binning = torch.empty(int(_L.svr_binning_state_size(R)) if R > 0 else 0, dtype=torch.uint8, device=dev)
    table = order_table_cache.get(octree_paths) if N > 0 else None
    _lib.check(_L.svr_rasterize_binning(W, H, N, _ptr(geomBuffer), _ptr(table), R, _ptr(binning), _ptr(img),
                                        SVR_FWD_PAIRS_READY if pairs_ready else 0, _stream(dev)),
               "rasterize_binning")
    return binning, img


# binding.cpp:26 -> forward.cu:792-879
def rasterize_voxels(n_samp_per_vox, image_width, image_height, tan_fovx, tan_fovy, cx, cy, w2c_matrix, c2w_matrix,
                     bg_color, need_depth, need_distortion, need_normal, track_max_w, octree_paths, vox_centers,
                     vox_lengths, geos, rgbs, geomBuffer, debug=False, num_rendered=None, pairs_ready=False,
                     binned=None):
    """num_rendered: R when the caller already has it (visible_voxels); otherwise it is read here.
    pairs_ready: visible_voxels ran with this octree's order table on geomBuffer.
    binned: (binningBuffer, imageBuffer) from rasterize_binning of the same R; only the compositor runs."""
    if vox_centers.ndimension() != 2 or vox_centers.size(1) != 3:
        raise RuntimeError("vox_centers must have dimensions (num_points, 3)")
    if rgbs.ndimension() != 2 or rgbs.size(1) != 3:
        raise RuntimeError("rgbs should be either (num_points, 3)")
    if vox_centers.size(0) != rgbs.size(0):
        raise RuntimeError("size mismatch")
    _require_cuda(octree_paths, vox_centers, vox_lengths, geos, rgbs, c2w_matrix)
    dev = vox_centers.device
    N = vox_centers.size(0)
    W, H = int(image_width), int(image_height)
    out_color = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    out_T = torch.empty((1, H, W), dtype=torch.float32, device=dev)
    empty = torch.empty(0, dtype=torch.float32, device=dev)
    ext = bool(need_depth or need_distortion or need_normal or track_max_w)
    out_depth = torch.empty((3, H, W), dtype=torch.float32, device=dev) if need_depth or need_distortion else empty
    out_normal = torch.empty((3, H, W), dtype=torch.float32, device=dev) if need_normal else empty
    max_w = torch.empty((N, 1), dtype=torch.float32, device=dev) if track_max_w else empty
    if num_rendered is None:
        R = ctypes.c_int64(0)
        if N > 0:
            _lib.check(_L.svr_num_rendered(N, _ptr(geomBuffer), _stream(dev), ctypes.byref(R)), "num_rendered")
        R = int(R.value)
    else:
        R = int(num_rendered)
    if binned is not None:
        binning, img = binned
        flags = SVR_FWD_BINNED
    else:
        img = torch.empty(int(_L.svr_image_state_size(W, H, R)), dtype=torch.uint8, device=dev)
        binning = torch.empty(int(_L.svr_binning_state_size(R)) if R > 0 else 0, dtype=torch.uint8, device=dev)
        flags = SVR_FWD_PAIRS_READY if pairs_ready else 0
    table = order_table_cache.get(octree_paths) if R > 0 else None
    centers, lengths, geos_c, rgbs_c = _f32(vox_centers), _f32(vox_lengths), _f32_a16(geos), _f32(rgbs)
    c2w = _f32(c2w_matrix)
    if ext:
        # forward.cu:262-394: depth / distortion / normal / max_w outputs
        _lib.check(_L.svr_rasterize_forward_ext(
            int(n_samp_per_vox), W, H, float(tan_fovx), float(tan_fovy), float(cx), float(cy), _ptr(c2w),
            float(bg_color), int(bool(need_depth)), int(bool(need_distortion)), int(bool(need_normal)),
            int(bool(track_max_w)), N, _ptr(centers), _ptr(lengths), _ptr(geos_c), _ptr(rgbs_c), _ptr(geomBuffer),
            _ptr(table), R, _ptr(binning), _ptr(img), _ptr(out_color), _ptr(out_T), _ptr(out_depth),
            _ptr(out_normal), _ptr(max_w), flags, _stream(dev)), "rasterize_voxels")
    else:
        _lib.check(_L.svr_rasterize_forward(int(n_samp_per_vox), W, H, float(tan_fovx), float(tan_fovy), float(cx),
                                            float(cy), _ptr(c2w), float(bg_color), N, _ptr(centers), _ptr(lengths),
                                            _ptr(geos_c), _ptr(rgbs_c), _ptr(geomBuffer), _ptr(table), R,
                                            _ptr(binning), _ptr(img), _ptr(out_color), _ptr(out_T), flags,
                                            _stream(dev)),
                   "rasterize_voxels")
    _debug(debug, dev, "rasterize_voxels")
    return R, binning, img, out_color, out_depth, out_normal, out_T, max_w


# binding.cpp:27 -> backward.cu:654-772
def rasterize_voxels_backward(R, n_samp_per_vox, image_width, image_height, tan_fovx, tan_fovy, cx, cy, w2c_matrix,
                              c2w_matrix, bg_color, octree_paths, vox_centers, vox_lengths, geos, rgbs, geomBuffer,
                              binningBuffer, imageBuffer, out_T, dL_dout_color, dL_dout_depth, dL_dout_normal,
                              dL_dout_T, lambda_R_concen, gt_color, lambda_ascending, lambda_dist, need_depth,
                              need_normal, out_D, out_N, debug=False):
    if vox_centers.ndimension() != 2 or vox_centers.size(1) != 3:
        raise RuntimeError("vox_centers must have dimensions (num_points, 3)")
    dev = vox_centers.device
    N = vox_centers.size(0)
    if N == 0:
        e = torch.empty(0, device=dev)
        return e, e, e
    W, H = int(image_width), int(image_height)
    dgeos = torch.empty((N, geos.size(1)), dtype=torch.float32, device=dev)
    drgbs = torch.empty((N, 3), dtype=torch.float32, device=dev)
    dsub = torch.empty((N, 1), dtype=torch.float32, device=dev)
    centers, lengths, geos_c, rgbs_c = _f32(vox_centers), _f32(vox_lengths), _f32_a16(geos), _f32(rgbs)
    dcol = _f32(dL_dout_color)
    dT = _f32(dL_dout_T) if dL_dout_T is not None else None  # the plain backward takes no T loss as NULL
    work = torch.empty(int(_L.svr_backward_workspace_size(int(R))), dtype=torch.uint8, device=dev)
    if need_depth or need_normal or lambda_R_concen > 0 or lambda_ascending > 0 or lambda_dist > 0:
        # backward.cu:160-510: depth / normal gradients and the regularizers
        def req(t, shape, what):
            if t is None or tuple(t.shape) != shape:
                raise RuntimeError(f"{what} must have shape {shape}")
            return _f32(t)
        dD = req(dL_dout_depth, (3, H, W), "dL_dout_depth") if need_depth else None
        dN = req(dL_dout_normal, (3, H, W), "dL_dout_normal") if need_normal else None
        gt = req(gt_color, (3, H, W), "gt_color") if lambda_R_concen > 0 else None
        oD = req(out_D, (3, H, W), "out_depth (forward run with lambda_dist > 0)") if lambda_dist > 0 else None
        _lib.check(_L.svr_rasterize_backward_ext(
            int(n_samp_per_vox), W, H, float(tan_fovx), float(tan_fovy), float(cx), float(cy),
            _ptr(_f32(c2w_matrix)), float(bg_color), N, _ptr(centers), _ptr(lengths), _ptr(geos_c), _ptr(rgbs_c),
            _ptr(geomBuffer), int(R), _ptr(binningBuffer), _ptr(imageBuffer), _ptr(_f32(out_T)), _ptr(dcol),
            _ptr(dD), _ptr(dN), _ptr(dT if dT is not None else torch.zeros((1, H, W), device=dev)),
            float(lambda_R_concen), _ptr(gt), float(lambda_ascending),
            float(lambda_dist), int(bool(need_depth)), int(bool(need_normal)), _ptr(oD), _ptr(work), _ptr(dgeos),
            _ptr(drgbs), _ptr(dsub), _stream(dev)), "rasterize_voxels_backward")
        _debug(debug, dev, "rasterize_voxels_backward")
        return dgeos, drgbs, dsub
    _lib.check(_L.svr_rasterize_backward(int(n_samp_per_vox), W, H, float(tan_fovx), float(tan_fovy), float(cx),
                                         float(cy), _ptr(_f32(c2w_matrix)), float(bg_color), N, _ptr(centers),
                                         _ptr(lengths), _ptr(geos_c), _ptr(rgbs_c), _ptr(geomBuffer), int(R),
                                         _ptr(binningBuffer), _ptr(imageBuffer), _ptr(_f32(out_T)), _ptr(dcol),
                                         _ptr(dT), _ptr(work), _ptr(dgeos), _ptr(drgbs), _ptr(dsub), _stream(dev)),
               "rasterize_voxels_backward")
    _debug(debug, dev, "rasterize_voxels_backward")
    return dgeos, drgbs, dsub


# binding.cpp:29 -> raster_state.cu:185-205
def unpack_ImageState(image_width, image_height, imageBuffer):
    W, H = int(image_width), int(image_height)
    gx, gy = tile_grid(W, H)
    dev = imageBuffer.device
    ranges = torch.zeros((gy, gx, 2), dtype=torch.int32, device=dev)
    tile_last = torch.zeros((gy, gx), dtype=torch.int32, device=dev)
    n_contrib = torch.zeros((H, W), dtype=torch.int32, device=dev)
    _lib.check(_L.svr_unpack_image_state(W, H, _ptr(imageBuffer), _ptr(ranges), _ptr(tile_last), _ptr(n_contrib),
                                         _stream(dev)), "unpack_ImageState")
    return ranges, tile_last, n_contrib


def unpack_geom_state(N, geomBuffer):
    """(n_duplicates, bbox [N,2], quadrant bitsets) of a preprocess call, as int32 bit patterns."""
    dev = geomBuffer.device
    n_dup = torch.zeros(N, dtype=torch.int32, device=dev)
    bbox = torch.zeros((N, 2), dtype=torch.int32, device=dev)
    qbits = torch.zeros(N, dtype=torch.int32, device=dev)
    if N > 0:
        _lib.check(_L.svr_unpack_geom_state(int(N), _ptr(geomBuffer), _ptr(n_dup), _ptr(bbox), _ptr(qbits),
                                            _stream(dev)), "unpack_geom_state")
    return n_dup, bbox, qbits


def export_sort_keys(image_width, image_height, R, binningBuffer, octree_paths):
    """Sorted keys in the reference's 128-bit form (hi, lo as int64 bit patterns) and values."""
    dev = binningBuffer.device
    hi = torch.zeros(R, dtype=torch.int64, device=dev)
    lo = torch.zeros(R, dtype=torch.int64, device=dev)
    vals = torch.zeros(R, dtype=torch.int32, device=dev)
    paths = octree_paths.reshape(-1).contiguous()
    if R > 0:
        _lib.check(_L.svr_export_sort_keys(int(image_width), int(image_height), int(R), _ptr(binningBuffer), None,
                                           _ptr(paths), paths.numel(), _ptr(hi), _ptr(lo), _ptr(vals),
                                           _stream(dev)), "export_sort_keys")
    return hi, lo, vals


# binding.cpp:36-37 -> sh_compute.cu:244-299
def sh_compute(D, indices, vox_centers, cam_pos, sh0, shs, visible=None):
    """visible: the geometry state whose in-frustum list `indices` is (visible_lists.geom_of), or None; with it
    the 16-coefficient layout takes the listed rows from the state (svr_sh_compute_visible, same rgbs)."""
    _require_cuda(vox_centers, cam_pos, sh0, shs)
    dev = vox_centers.device
    N = vox_centers.size(0)
    M = 1 + shs.size(1)
    rgbs = torch.empty((N, 3), dtype=torch.float32, device=dev)
    if visible is not None and M == 16 and N > 0:
        _lib.check(_L.svr_sh_compute_visible(int(D), N, _ptr(visible), _ptr(_f32(vox_centers)), _ptr(_f32(cam_pos)),
                                             _ptr(_f32_a16(sh0)), _ptr(_f32_a16(shs)), M, _ptr(rgbs), _stream(dev)),
                   "sh_compute")
        return rgbs
    idx = indices.to(device=dev, dtype=torch.int64).contiguous() if indices is not None else None
    n_idx = idx.numel() if idx is not None else 0
    _lib.check(_L.svr_sh_compute(int(D), n_idx, _ptr(idx), N, _ptr(_f32(vox_centers)), _ptr(_f32(cam_pos)),
                                 _ptr(_f32(sh0)), _ptr(_f32(shs)), M, _ptr(rgbs), _stream(dev)), "sh_compute")
    return rgbs


def sh_compute_bw(D, M, indices, vox_centers, cam_pos, rgbs, dL_drgbs, visible=None):
    """visible: as for sh_compute (svr_sh_compute_bw_visible, same gradients)."""
    dev = vox_centers.device
    N = vox_centers.size(0)
    dsh0 = torch.empty((N, 3), dtype=torch.float32, device=dev)
    dshs = torch.empty((N, M - 1, 3), dtype=torch.float32, device=dev)
    if visible is not None and M == 16 and N > 0:
        _lib.check(_L.svr_sh_compute_bw_visible(int(D), int(M), N, _ptr(visible), _ptr(_f32(vox_centers)),
                                                _ptr(_f32(cam_pos)), _ptr(_f32(dL_drgbs)), _ptr(dsh0), _ptr(dshs),
                                                _stream(dev)), "sh_compute_bw")
        return dsh0, dshs
    idx = indices.to(device=dev, dtype=torch.int64).contiguous() if indices is not None else None
    n_idx = idx.numel() if idx is not None else 0
    _lib.check(_L.svr_sh_compute_bw(int(D), int(M), n_idx, _ptr(idx), N, _ptr(_f32(vox_centers)),
                                    _ptr(_f32(cam_pos)), _ptr(_f32(dL_drgbs)), _ptr(dsh0), _ptr(dshs),
                                    _stream(dev)), "sh_compute_bw")
    return dsh0, dshs


def _check_adam(param, grad, exp_avg, exp_avg_sq):
    for t in (param, grad, exp_avg, exp_avg_sq):
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise RuntimeError("adam step expects contiguous float32 tensors")
    _require_cuda(param, grad, exp_avg, exp_avg_sq)


# binding.cpp:48-49 -> adam_step.cu:106-162
def unbiased_adam_step(sparse, param, grad, exp_avg, exp_avg_sq, step, lr, beta1, beta2, eps):
    _check_adam(param, grad, exp_avg, exp_avg_sq)
    _lib.check(_L.svr_unbiased_adam_step(int(bool(sparse)), param.numel(), _ptr(param), _ptr(grad), _ptr(exp_avg),
                                         _ptr(exp_avg_sq), float(step), float(lr), float(beta1), float(beta2),
                                         float(eps), _stream(param.device)), "unbiased_adam_step")


def adam_step_multi(biased, sparse, items):
    """Several tensors' Adam updates in one launch (svr_adam_step_multi; no reference counterpart, the reference
    calls unbiased_adam_step / biased_adam_step once per tensor): items is a list of (param, grad, exp_avg,
    exp_avg_sq, step, lr, beta1, beta2, eps) with every tensor on one device; each update is bit-equal to the
    per-tensor call."""
    if not items:
        return
    dev = items[0][0].device
    for it in items:
        _check_adam(*it[:4])
        if it[0].device != dev:
            raise RuntimeError("adam_step_multi: tensors on different devices")
        if not (it[0].numel() == it[1].numel() == it[2].numel() == it[3].numel()):
            raise RuntimeError("adam_step_multi: param, grad and moments differ in size")
    n = len(items)
    arr = lambda ct, vals: (ct * n)(*vals)  # noqa: E731
    _lib.check(_L.svr_adam_step_multi(
        int(bool(biased)), int(bool(sparse)), n,
        arr(ctypes.c_int64, [it[0].numel() for it in items]),
        arr(ctypes.c_void_p, [_ptr(it[0]) for it in items]),
        arr(ctypes.c_void_p, [_ptr(it[1]) for it in items]),
        arr(ctypes.c_void_p, [_ptr(it[2]) for it in items]),
        arr(ctypes.c_void_p, [_ptr(it[3]) for it in items]),
        arr(ctypes.c_double, [float(it[4]) for it in items]),
        arr(ctypes.c_double, [float(it[5]) for it in items]),
        arr(ctypes.c_double, [float(it[6]) for it in items]),
        arr(ctypes.c_double, [float(it[7]) for it in items]),
        arr(ctypes.c_float, [float(it[8]) for it in items]),
        _stream(dev)), "adam_step_multi")


def adam_bias_table(first_step, count, lr, beta1, beta2, device):
    """Device table of the host-form bias terms (svr_adam_bias_terms) of steps first_step .. first_step + count - 1:
    int32 [first_step, count] followed by count (step_size, rb2) float32 pairs, in one int32 tensor."""
    terms = np.empty((count, 2), dtype=np.float32)
    a, b = ctypes.c_float(0), ctypes.c_float(0)
    for i in range(count):
        _lib.check(_L.svr_adam_bias_terms(float(first_step + i), float(lr), float(beta1), float(beta2),
                                          ctypes.byref(a), ctypes.byref(b)), "adam_bias_terms")
        terms[i] = (a.value, b.value)
    head = np.array([first_step, count], dtype=np.int32)
    return torch.from_numpy(np.concatenate([head, terms.reshape(-1).view(np.int32)])).to(device)


def unbiased_adam_step_device(sparse, param, grad, exp_avg, exp_avg_sq, step_dev, coef_dev, lr, beta1, beta2, eps,
                              bias_table=None, skip_dev=None):
    """Capture-safe form (no reference counterpart): the step count lives in step_dev (int32 [1]) and is
    incremented on the device; the bias terms of the new count come from bias_table (adam_bias_table; pow() on
    the device outside it); coef_dev (float32 [3]) is scratch.  skip_dev (int32 [1], device): nonzero leaves
    everything unchanged."""
    _check_adam(param, grad, exp_avg, exp_avg_sq)
    if coef_dev.numel() < 3:
        raise RuntimeError("coef_dev needs 3 floats")
    _lib.check(_L.svr_unbiased_adam_step_device(int(bool(sparse)), param.numel(), _ptr(param), _ptr(grad),
                                                _ptr(exp_avg), _ptr(exp_avg_sq), _ptr(step_dev), _ptr(bias_table),
                                                _ptr(skip_dev), _ptr(coef_dev), float(lr), float(beta1),
                                                float(beta2), float(eps), _stream(param.device)),
               "unbiased_adam_step_device")


def biased_adam_step(sparse, param, grad, exp_avg, exp_avg_sq, lr, beta1, beta2, eps):
    _check_adam(param, grad, exp_avg, exp_avg_sq)
    _lib.check(_L.svr_biased_adam_step(int(bool(sparse)), param.numel(), _ptr(param), _ptr(grad), _ptr(exp_avg),
                                       _ptr(exp_avg_sq), float(lr), float(beta1), float(beta2), float(eps),
                                       _stream(param.device)), "biased_adam_step")


# binding.cpp:44-46 -> utils.cu:251-298
def voxel_order_rank(octree_paths):
    paths = octree_paths.reshape(-1).contiguous()
    out = torch.empty((8, paths.numel()), dtype=torch.int64, device=paths.device)
    _lib.check(_L.svr_voxel_order_rank(paths.numel(), _ptr(paths), _ptr(out), _stream(paths.device)),
               "voxel_order_rank")
    return out


def ijk_2_octpath(ijk, octlevel):
    ijk = ijk.contiguous()
    lv = octlevel.reshape(-1).contiguous()
    out = torch.empty((lv.numel(), 1), dtype=torch.int64, device=ijk.device)
    _lib.check(_L.svr_ijk_2_octpath(lv.numel(), _ptr(ijk), _ptr(lv), _ptr(out), _stream(ijk.device)),
               "ijk_2_octpath")
    return out


def octpath_2_ijk(octpath, octlevel):
    paths = octpath.reshape(-1).contiguous()
    lv = octlevel.reshape(-1).contiguous()
    out = torch.empty((lv.numel(), 3), dtype=torch.int64, device=paths.device)
    _lib.check(_L.svr_octpath_2_ijk(lv.numel(), _ptr(paths), _ptr(lv), _ptr(out), _stream(paths.device)),
               "octpath_2_ijk")
    return out


# ---------------------------------------------------------------------------
# binding.cpp:31-34 -> geo_params_gather.cu

def _i64(t):
    return t.contiguous() if t.dtype == torch.int64 else t.long().contiguous()


def gather_triinterp_geo_params(vox_key, care_idx, grid_pts):
    _require_cuda(vox_key, care_idx, grid_pts)
    N = vox_key.size(0)
    geo_params = torch.empty((N, 8), dtype=grid_pts.dtype, device=grid_pts.device)
    key, idx, pts = _i64(vox_key), _i64(care_idx), _f32(grid_pts)
    _lib.check(_L.svr_gather_geo_params(idx.numel(), _ptr(key), _ptr(idx), _ptr(pts), _ptr(geo_params),
                                        _stream(grid_pts.device)), "gather_triinterp_geo_params")
    return geo_params


def gather_triinterp_geo_params_bw(vox_key, care_idx, num_grid_pts, dL_dgeo_params):
    _require_cuda(vox_key, care_idx, dL_dgeo_params)
    dL_dgrid_pts = torch.zeros((int(num_grid_pts), 1), dtype=dL_dgeo_params.dtype, device=dL_dgeo_params.device)
    key, idx, g = _i64(vox_key), _i64(care_idx), _f32(dL_dgeo_params)
    _lib.check(_L.svr_gather_geo_params_bw(idx.numel(), _ptr(key), _ptr(idx), _ptr(g), _ptr(dL_dgrid_pts),
                                           _stream(g.device)), "gather_triinterp_geo_params_bw")
    return dL_dgrid_pts


def gather_triinterp_feat_params(vox_key, care_idx, grid_pts):
    _require_cuda(vox_key, care_idx, grid_pts)
    N, n_dim = vox_key.size(0), grid_pts.size(1)
    if n_dim != 3:
        raise RuntimeError("Only support n_dim=3 now.")
    feat = torch.empty((N, 8, n_dim), dtype=grid_pts.dtype, device=grid_pts.device)
    key, idx, pts = _i64(vox_key), _i64(care_idx), _f32(grid_pts)
    _lib.check(_L.svr_gather_feat_params(idx.numel(), n_dim, _ptr(key), _ptr(idx), _ptr(pts), _ptr(feat),
                                         _stream(grid_pts.device)), "gather_triinterp_feat_params")
    return feat


def gather_triinterp_feat_params_bw(vox_key, care_idx, num_grid_pts, dL_dfeat_params):
    _require_cuda(vox_key, care_idx, dL_dfeat_params)
    n_dim = dL_dfeat_params.size(2)
    if n_dim != 3:
        raise RuntimeError("Only support n_dim=3")
    dL_dgrid_pts = torch.zeros((int(num_grid_pts), n_dim), dtype=dL_dfeat_params.dtype,
                               device=dL_dfeat_params.device)
    key, idx, g = _i64(vox_key), _i64(care_idx), _f32(dL_dfeat_params)
    _lib.check(_L.svr_gather_feat_params_bw(idx.numel(), n_dim, _ptr(key), _ptr(idx), _ptr(g), _ptr(dL_dgrid_pts),
                                            _stream(g.device)), "gather_triinterp_feat_params_bw")
    return dL_dgrid_pts


# binding.cpp:39 -> tv_compute.cu:83-111 (in place on grid_pts_grad)
def total_variation_bw(grid_pts, vox_key, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad):
    _require_cuda(grid_pts, vox_key, vox_size_inv, grid_pts_grad)
    if not grid_pts_grad.is_contiguous() or grid_pts_grad.dtype != torch.float32:
        raise RuntimeError("grid_pts_grad must be a contiguous float32 tensor (updated in place)")
    N, C = vox_key.size(0), grid_pts.size(1)
    pts, key = _f32(grid_pts), _i64(vox_key)
    vsi = _f32(vox_size_inv) if vox_size_inv is not None else None
    ws = torch.empty(int(_L.svr_total_variation_workspace_size(N, C)) if tv_sparse else 0, dtype=torch.uint8,
                     device=grid_pts.device)
    _lib.check(_L.svr_total_variation_bw(N, C, _ptr(pts), _ptr(key), float(weight), _ptr(vsi), int(bool(no_tv_s)),
                                         int(bool(tv_sparse)), _ptr(ws), _ptr(grid_pts_grad),
                                         _stream(grid_pts.device)), "total_variation_bw")


# binding.cpp:28, 41-43 -> raster_state.cu:227-260, utils.cu:188-245
def filter_geomState(ori_P, indices, geomState):
    _require_cuda(indices, geomState)
    P = indices.numel()
    new = torch.empty(int(_L.svr_geom_state_size(P)) if P > 0 else 0, dtype=torch.uint8, device=geomState.device)
    if P > 0:
        idx = _i64(indices)
        _lib.check(_L.svr_filter_geom_state(int(ori_P), P, _ptr(idx), _ptr(geomState), _ptr(new),
                                            _stream(geomState.device)), "filter_geomState")
    return new


def is_in_cone(tanfovx, tanfovy, near, w2c_matrix, pts):
    _require_cuda(w2c_matrix, pts)
    N = pts.size(0)
    is_in = torch.zeros(N, dtype=torch.bool, device=pts.device)
    if N > 0:
        p, w2c = _f32(pts), _f32(w2c_matrix)
        _lib.check(_L.svr_is_in_cone(N, float(tanfovx), float(tanfovy), float(near), _ptr(w2c), _ptr(p),
                                     _ptr(is_in), _stream(pts.device)), "is_in_cone")
    return is_in


def compute_rd(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix):
    _require_cuda(c2w_matrix)
    rd = torch.empty((3, int(height), int(width)), dtype=torch.float32, device=c2w_matrix.device)
    _lib.check(_L.svr_compute_rd(int(width), int(height), float(cx), float(cy), float(tanfovx), float(tanfovy),
                                 _ptr(_f32(c2w_matrix)), _ptr(rd), _stream(c2w_matrix.device)), "compute_rd")
    return rd


def depth2pts(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix, depth):
    _require_cuda(c2w_matrix, depth)
    pts = torch.empty((3, int(height), int(width)), dtype=torch.float32, device=c2w_matrix.device)
    _lib.check(_L.svr_depth2pts(int(width), int(height), float(cx), float(cy), float(tanfovx), float(tanfovy),
                                _ptr(_f32(c2w_matrix)), _ptr(_f32(depth)), _ptr(pts), _stream(c2w_matrix.device)),
               "depth2pts")
    return pts

