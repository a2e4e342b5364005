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
    raise NotImplementedError("is_in_cone is outside the rasterize path implemented by sv_raster_amd")


def compute_rd(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix):
    raise NotImplementedError("compute_rd is outside the rasterize path implemented by sv_raster_amd")


def depth2pts(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix, depth):
    raise NotImplementedError("depth2pts is outside the rasterize path implemented by sv_raster_amd")
