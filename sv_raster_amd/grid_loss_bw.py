"""Total-variation gradient on grid-point parameters
(backends/new_cuda/new_svraster_cuda/grid_loss_bw.py:13-17, tv_compute.cu:23-111).

`total_variation` adds, in place into grid_pts_grad [M, C], the gradient of
the L1 differences of grid-point values along the 12 edges of every voxel's
corner cube, scaled by weight (times 0.01 / voxel size unless no_tv_s).  With
tv_sparse, only (voxel, channel) items with a non-zero corner gradient take
part, judged on grid_pts_grad as it stands before the call.
"""
from . import _C


def total_variation(grid_pts, vox_key, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad):
    assert grid_pts.shape == grid_pts_grad.shape
    assert len(vox_key.shape) == 2 and vox_key.shape[1] == 8
    assert vox_key.shape[0] == vox_size_inv.numel()
    _C.total_variation_bw(grid_pts, vox_key, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad)
