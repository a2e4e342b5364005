"""Drop-in rasterizer entry point: the same RasterSettings fields, argument
layout, validation messages and autograd plumbing as the reference backend
(backends/new_cuda/new_svraster_cuda/renderer.py), running on the HIP kernels
of libsvraster_hip.so.
"""
from typing import NamedTuple

import torch

from . import _C


class RasterSettings(NamedTuple):
    """renderer.py:15-36 of the reference (same fields, order and defaults)."""
    color_mode: str
    n_samp_per_vox: int
    image_width: int
    image_height: int
    tanfovx: float
    tanfovy: float
    cx: float
    cy: float
    w2c_matrix: torch.Tensor
    c2w_matrix: torch.Tensor
    bg_color: float = 0
    near: float = 0.02
    need_depth: bool = False
    need_normal: bool = False
    track_max_w: bool = False
    lambda_R_concen: float = 0
    lambda_ascending: float = 0
    lambda_dist: float = 0
    gt_color: torch.Tensor = torch.empty(0)
    debug: bool = False


def rasterize_voxels(
        raster_settings: RasterSettings,
        octree_paths: torch.Tensor,
        vox_centers: torch.Tensor,
        vox_lengths: torch.Tensor,
        vox_fn,
):
    """renderer.py:39-127: validate, preprocess, gather per-voxel parameters
    through vox_fn for the in-frustum voxels, then the differentiable raster.
    Returns (color [3,H,W], depth, normal, T [1,H,W], max_w)."""
    if not isinstance(raster_settings, RasterSettings):
        raise Exception("Expect RasterSettings as first argument.")
    if raster_settings.n_samp_per_vox > _C.MAX_N_SAMP or raster_settings.n_samp_per_vox < 1:
        raise Exception(f"n_samp_per_vox should be in range [1, {_C.MAX_N_SAMP}].")

    N = octree_paths.numel()
    device = octree_paths.device
    if vox_centers.shape[0] != N or vox_lengths.numel() != N:
        raise Exception("Size mismatched.")
    if len(vox_centers.shape) != 2 or vox_centers.shape[1] != 3:
        raise Exception("Expect vox_centers in shape [N, 3].")
    if raster_settings.w2c_matrix.device != device or \
            raster_settings.c2w_matrix.device != device or \
            vox_centers.device != device or \
            vox_lengths.device != device:
        raise Exception("Device mismatch.")

    # the camera position handed to vox_fn (renderer.py:85 passes the view c2w[:3, 3]); made contiguous here,
    # before the host synchronisation below, so that the colour evaluation does not copy it on the critical path
    cam_pos = raster_settings.c2w_matrix[:3, 3].contiguous()

    n_duplicates, geomBuffer = _C.rasterize_preprocess(
        raster_settings.image_width,
        raster_settings.image_height,
        raster_settings.tanfovx,
        raster_settings.tanfovy,
        raster_settings.cx,
        raster_settings.cy,
        raster_settings.w2c_matrix,
        raster_settings.c2w_matrix,
        raster_settings.near,
        octree_paths,
        vox_centers,
        vox_lengths,
        raster_settings.debug,
    )
    # torch.where(n_duplicates > 0)[0] and the entry count R, read with one host synchronisation; the part of
    # the binning that does not need R is queued behind it, and the rest right after it, before vox_fn, so the
    # device bins while the host evaluates the colours (binning needs the geometry only)
    table = _C.order_table_cache.get(octree_paths) if N > 0 else None
    in_frusts_idx, num_rendered = _C.visible_voxels(geomBuffer, N, device, table)
    binned = _C.rasterize_binning(raster_settings.image_width, raster_settings.image_height, octree_paths,
                                  geomBuffer, num_rendered, pairs_ready=table is not None)

    vox_params = vox_fn(in_frusts_idx, cam_pos, raster_settings.color_mode)
    geos = vox_params['geos']
    rgbs = vox_params['rgbs']
    subdiv_p = vox_params['subdiv_p']

    if geos.shape != (N, 8):
        raise Exception(f"Expect geos in ({N}, 8) but got", geos.shape)
    if rgbs.shape[0] != N:
        raise Exception(f"Expect rgbs in ({N}, 3) but got", rgbs.shape)
    if subdiv_p.shape[0] != N:
        raise Exception(f"Expect subdiv_p in ({N}, 1) but got", subdiv_p.shape)

    if geos.device != device:
        raise Exception("Device mismatch: geos.")
    if rgbs.device != device:
        raise Exception("Device mismatch: rgbs.")
    if subdiv_p.device != device:
        raise Exception("Device mismatch: subdiv_p.")

    if raster_settings.lambda_R_concen > 0:
        if len(raster_settings.gt_color.shape) != 3 or \
                raster_settings.gt_color.shape[0] != 3 or \
                raster_settings.gt_color.shape[1] != raster_settings.image_height or \
                raster_settings.gt_color.shape[2] != raster_settings.image_width:
            raise Exception("Except gt_color in shape of [3, H, W]")
        if raster_settings.gt_color.device != device:
            raise Exception("Device mismatch.")

    return _RasterizeVoxels.apply(
        raster_settings,
        geomBuffer,
        octree_paths,
        vox_centers,
        vox_lengths,
        geos,
        rgbs,
        subdiv_p,
        num_rendered,
        binned,
    )


class _RasterizeVoxels(torch.autograd.Function):
    """renderer.py:130-248: forward saves the opaque raster state; backward
    returns gradients for geos, rgbs and the subdivision priority."""

    @staticmethod
    def forward(ctx, raster_settings, geomBuffer, octree_paths, vox_centers, vox_lengths, geos, rgbs, subdiv_p,
                num_rendered=None, binned=None):
        need_distortion = raster_settings.lambda_dist > 0
        num_rendered, binningBuffer, imgBuffer, out_color, out_depth, out_normal, out_T, max_w = \
            _C.rasterize_voxels(
                raster_settings.n_samp_per_vox,
                raster_settings.image_width,
                raster_settings.image_height,
                raster_settings.tanfovx,
                raster_settings.tanfovy,
                raster_settings.cx,
                raster_settings.cy,
                raster_settings.w2c_matrix,
                raster_settings.c2w_matrix,
                raster_settings.bg_color,
                raster_settings.need_depth,
                need_distortion,
                raster_settings.need_normal,
                raster_settings.track_max_w,
                octree_paths,
                vox_centers,
                vox_lengths,
                geos,
                rgbs,
                geomBuffer,
                raster_settings.debug,
                num_rendered=num_rendered,
                binned=binned,
            )
        ctx.raster_settings = raster_settings
        ctx.num_rendered = num_rendered
        # outputs the loss does not use arrive as None (no zero tensors materialised): the backward takes a
        # missing T gradient as none at all
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(
            octree_paths, vox_centers, vox_lengths,
            geos, rgbs,
            geomBuffer, binningBuffer, imgBuffer, out_T, out_depth, out_normal)
        ctx.mark_non_differentiable(max_w)
        return out_color, out_depth, out_normal, out_T, max_w

    @staticmethod
    def backward(ctx, dL_dout_color, dL_dout_depth, dL_dout_normal, dL_dout_T, dL_dmax_w):
        raster_settings = ctx.raster_settings
        num_rendered = ctx.num_rendered
        octree_paths, vox_centers, vox_lengths, \
            geos, rgbs, \
            geomBuffer, binningBuffer, imgBuffer, out_T, out_depth, out_normal = ctx.saved_tensors
        shape = (3, raster_settings.image_height, raster_settings.image_width)
        if dL_dout_color is None:
            dL_dout_color = torch.zeros(shape, device=geos.device)
        if raster_settings.need_depth and dL_dout_depth is None:
            dL_dout_depth = torch.zeros(shape, device=geos.device)
        if raster_settings.need_normal and dL_dout_normal is None:
            dL_dout_normal = torch.zeros(shape, device=geos.device)

        dL_dgeos, dL_drgbs, subdiv_p_bw = _C.rasterize_voxels_backward(
            num_rendered,
            raster_settings.n_samp_per_vox,
            raster_settings.image_width,
            raster_settings.image_height,
            raster_settings.tanfovx,
            raster_settings.tanfovy,
            raster_settings.cx,
            raster_settings.cy,
            raster_settings.w2c_matrix,
            raster_settings.c2w_matrix,
            raster_settings.bg_color,
            octree_paths,
            vox_centers,
            vox_lengths,
            geos,
            rgbs,
            geomBuffer,
            binningBuffer,
            imgBuffer,
            out_T,
            dL_dout_color,
            dL_dout_depth,
            dL_dout_normal,
            dL_dout_T,
            raster_settings.lambda_R_concen,
            raster_settings.gt_color,
            raster_settings.lambda_ascending,
            raster_settings.lambda_dist,
            raster_settings.need_depth,
            raster_settings.need_normal,
            out_depth,
            out_normal,
            raster_settings.debug,
        )
        return None, None, None, None, None, dL_dgeos, dL_drgbs, subdiv_p_bw, None, None


def _no_grad_input(*tensors):
    """SH_eval differentiates the coefficients only (renderer.py:265-270 raise NotImplementedError for the
    geometry inputs)."""
    if any(torch.is_tensor(t) and t.requires_grad for t in tensors):
        raise NotImplementedError


class SH_eval(torch.autograd.Function):
    """renderer.py:251-316: rgbs [N, 3] = SH(active_sh_degree) of the direction cam_pos -> voxel centre (or of
    `viewdir` [N, 3] when given), + 0.5, clamped at 0; rows outside `idx` are zero (idx None: all voxels).
    Gradients flow to sh0 and shs."""

    @staticmethod
    def forward(ctx, active_sh_degree, idx, vox_centers, cam_pos, viewdir, sh0, shs):
        _no_grad_input(vox_centers, cam_pos, viewdir)
        # a given view direction is evaluated as the direction from the origin to it
        points, origin = (vox_centers, cam_pos) if viewdir is None else (viewdir, torch.zeros_like(cam_pos))
        sel = torch.empty(0, dtype=torch.int64) if idx is None else idx
        # an in-frustum list from rasterize_voxels: its rows come from the geometry state (no marking pass)
        visible = _C.visible_lists.geom_of(idx, points.shape[0])
        rgbs = _C.sh_compute(active_sh_degree, sel, points, origin, sh0, shs, visible=visible)
        ctx.active_sh_degree, ctx.M, ctx.visible = active_sh_degree, shs.shape[1] + 1, visible
        ctx.save_for_backward(sel, points, origin, rgbs)
        return rgbs

    @staticmethod
    def backward(ctx, dL_drgbs):
        sel, points, origin, rgbs = ctx.saved_tensors
        grads = _C.sh_compute_bw(ctx.active_sh_degree, ctx.M, sel, points, origin, rgbs, dL_drgbs,
                                 visible=ctx.visible)
        return (None,) * 5 + tuple(grads)


def _check_gather(vox_key, care_idx):
    assert vox_key.dim() == 2 and vox_key.shape[1] == 8
    assert care_idx.dim() == 1


class GatherGeoParams(torch.autograd.Function):
    """renderer.py:319-345: geo_params[idx, k] = grid_pts[vox_key[idx, k]] ([N, 8]) for idx in care_idx;
    the backward scatters the rows back onto the grid points."""

    @staticmethod
    def forward(ctx, vox_key, care_idx, grid_pts):
        _check_gather(vox_key, care_idx)
        assert grid_pts.numel() == grid_pts.shape[0]  # one value per grid point
        ctx.grid_shape = grid_pts.shape
        ctx.save_for_backward(vox_key, care_idx)
        return _C.gather_triinterp_geo_params(vox_key, care_idx, grid_pts)

    @staticmethod
    def backward(ctx, dL_dgeo_params):
        vox_key, care_idx = ctx.saved_tensors
        n_pts = ctx.grid_shape[0]
        dL_dgrid = _C.gather_triinterp_geo_params_bw(vox_key, care_idx, n_pts, dL_dgeo_params)
        # shaped like the grid ([M, 1] in the reference's models; a flat [M] grid also works)
        return None, None, dL_dgrid.view(ctx.grid_shape)


class GatherFeatParams(torch.autograd.Function):
    """renderer.py:348-374: feat_params[idx, k] = grid_pts[vox_key[idx, k]] ([N, 8, C]) for idx in care_idx."""

    @staticmethod
    def forward(ctx, vox_key, care_idx, grid_pts):
        _check_gather(vox_key, care_idx)
        assert grid_pts.dim() == 2
        ctx.n_pts = grid_pts.shape[0]
        ctx.save_for_backward(vox_key, care_idx)
        return _C.gather_triinterp_feat_params(vox_key, care_idx, grid_pts)

    @staticmethod
    def backward(ctx, dL_dfeat_params):
        vox_key, care_idx = ctx.saved_tensors
        return None, None, _C.gather_triinterp_feat_params_bw(vox_key, care_idx, ctx.n_pts, dL_dfeat_params)


def mark_n_duplicates(image_width, image_height, tanfovx, tanfovy, cx, cy, w2c_matrix, c2w_matrix, near,
                      octree_paths, vox_centers, vox_lengths, return_buffer=False, debug=False):
    """renderer.py:377-405: the preprocess alone; n_duplicates [N] (int32), and the geometry state when
    return_buffer is set."""
    out = _C.rasterize_preprocess(image_width, image_height, tanfovx, tanfovy, cx, cy, w2c_matrix, c2w_matrix,
                                  near, octree_paths, vox_centers, vox_lengths, debug)
    return out if return_buffer else out[0]


def mark_max_samp_rate(cameras, octree_paths, vox_centers, vox_lengths, near=0.02):
    """renderer.py:408-423: per voxel, the largest length / (camera-axis distance x pixel size) over the
    cameras that see it beyond `near` ([N] float32, 0 if none).  All cameras in one batch pass
    (batch.mark_max_samp_rate) instead of a preprocess per camera."""
    from . import batch
    return batch.mark_max_samp_rate(cameras, octree_paths, vox_centers, vox_lengths, near=near)


def mark_near(cameras, octree_paths, vox_centers, vox_lengths, near=0.2):
    """renderer.py:426-438: True for voxels that some camera sees within `near` + the voxel's length along its
    axis ([N] bool).  All cameras in one batch pass (batch.mark_near)."""
    from . import batch
    return batch.mark_near(cameras, octree_paths, vox_centers, vox_lengths, near=near)
