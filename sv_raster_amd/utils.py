"""Octree utilities (backends/new_cuda/new_svraster_cuda/utils.py:13-58)."""
import torch

from . import _C


def voxel_order_rank(octree_paths):
    """The eight order ranks octpath ^ repeat(q) of every voxel, [8, N] (utils.py:16-19)."""
    return _C.voxel_order_rank(octree_paths)


def ijk_2_octpath(ijk, octlevel):
    """utils.py:47-54."""
    assert torch.is_tensor(ijk) and torch.is_tensor(octlevel)
    assert len(ijk.shape) == 2 and ijk.shape[1] == 3
    assert ijk.numel() == octlevel.numel() * 3
    assert ijk.dtype == torch.int64
    assert octlevel.dtype == torch.int8
    return _C.ijk_2_octpath(ijk, octlevel)


def octpath_2_ijk(octpath, octlevel):
    """utils.py:57-62."""
    assert torch.is_tensor(octpath) and torch.is_tensor(octlevel)
    assert octpath.numel() == octlevel.numel()
    assert octpath.dtype == torch.int64
    assert octlevel.dtype == torch.int8
    return _C.octpath_2_ijk(octpath, octlevel)


def is_in_cone(pts, cam):
    """utils.py:19-29: bool [N], point in front of cam.near and inside the field-of-view cone."""
    assert torch.is_tensor(pts)
    assert pts.device == cam.w2c.device
    assert len(pts.shape) == 2
    assert pts.shape[1] == 3
    return _C.is_in_cone(cam.tanfovx, cam.tanfovy, cam.near, cam.w2c, pts)


def compute_rd(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix):
    """utils.py:32-34: per-pixel world ray directions [3, H, W] (unnormalised)."""
    assert torch.is_tensor(c2w_matrix)
    return _C.compute_rd(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix)


def depth2pts(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix, depth):
    """utils.py:37-41: world points [3, H, W] at the given per-pixel depth along the rays."""
    assert torch.is_tensor(c2w_matrix)
    assert depth.device == c2w_matrix.device
    assert depth.numel() == width * height
    return _C.depth2pts(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix, depth)
