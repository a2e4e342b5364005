"""Trained sparse-voxel models in the reference's checkpoint format, turned
into the rasterizer's inputs.

Format (SRC/sparse_voxel_gears/io.py:17-46, 49-96): a `torch.save`d dict with
`backend_name`, `color_layout` ('voxel' here), `geo_layout` ('trilinear'),
`active_sh_degree`, `scene_center` [3], `inside_extent` [1], `scene_extent`
[1], `octpath` [N, 1] int64, `octlevel` [N, 1] int8, `_geo_grid_pts` [M, 1],
`_sh0` [N, 3], `_shs` [N, K, 3] and `quantized`.  A quantized file stores
`_geo_grid_pts` as one {index uint8, codebook [256]} dict and `_sh0` / `_shs`
as lists of such dicts, one per channel slice along dim 1 (io.py:118-128);
the value of an element is codebook[index] (io.py:176-177).

Reading never executes code from the file: `torch.load(weights_only=True)`
(the format holds only tensors, strings, numbers and lists / dicts of them).

The voxel geometry comes from the HIP octree codec: centres
scene_min + (ijk + 0.5) * size with size = scene_extent * 2^-level
(octree_utils.py:59-68, 107-134), and each voxel's eight corner grid points
are the distinct finest-level corner coordinates (octree_utils.py:137-154,
213-232), so `VoxelModel.render` gathers corner densities with
`GatherGeoParams` and colours with `SH_eval` exactly as the reference's
`SVRenderer.vox_fn` / `render` do (SRC/sparse_voxel_gears/renderer.py:
59-153, 155-230) for the voxel colour layout.
"""
import os

import torch

from . import meta, renderer, utils

FORMAT_BACKEND = "new_cuda"
_SUBTREE_SHIFT = torch.tensor([[(i >> 2) & 1, (i >> 1) & 1, i & 1] for i in range(8)], dtype=torch.int64)


def dequantize(q):
    """codebook[index] of one quantized block (io.py:176-177)."""
    return q["codebook"][q["index"].long()]


def _dequantize_state(sd):
    sd["_geo_grid_pts"] = dequantize(sd["_geo_grid_pts"])
    sd["_sh0"] = torch.cat([dequantize(q) for q in sd["_sh0"]], dim=1)
    sd["_shs"] = torch.cat([dequantize(q) for q in sd["_shs"]], dim=1)


def read_state(path):
    """The raw state dict of a checkpoint (quantized blocks expanded), on the CPU."""
    sd = torch.load(path, map_location="cpu", weights_only=True)
    backend = sd.get("backend_name")
    if backend not in (None, FORMAT_BACKEND):
        raise RuntimeError(f"Checkpoint backend {backend!r} does not match model backend {FORMAT_BACKEND!r}.")
    if sd.get("color_layout", "voxel") != "voxel" or sd.get("geo_layout", "trilinear") != "trilinear":
        raise RuntimeError("This checkpoint does not use the per-voxel colour / trilinear density layout of the "
                           "new_cuda rasterizer.")
    if sd.get("quantized", False):
        _dequantize_state(sd)
    return sd


def voxel_geometry(octpath, octlevel, scene_center, scene_extent):
    """Voxel centres [N, 3] and sizes [N, 1] (float32) of octree paths / levels (octree_utils.py:107-134)."""
    octpath = octpath.reshape(-1, 1)
    octlevel = octlevel.reshape(-1, 1).to(torch.int8)
    scene_min = scene_center - 0.5 * scene_extent
    size = torch.ldexp(scene_extent, -octlevel.long())
    ijk = utils.octpath_2_ijk(octpath, octlevel)
    return scene_min + (ijk + 0.5) * size, size


def grid_pts_link(octpath, octlevel):
    """(grid_pts_key [M, 3] int64, vox_key [N, 8] int64): the distinct finest-level integer coordinates of
    the voxels' corners, in lexicographic order, and each voxel's eight indices into them
    (octree_utils.py:137-154, 213-232; corner c = (c >> 2 & 1, c >> 1 & 1, c & 1))."""
    octlevel = octlevel.reshape(-1, 1).to(torch.int8)
    ijk = utils.octpath_2_ijk(octpath.reshape(-1, 1), octlevel)
    shift = (meta.MAX_NUM_LEVELS - octlevel.long()).view(-1, 1, 1)
    corners = (ijk.view(-1, 1, 3) << shift) + (_SUBTREE_SHIFT.to(ijk.device).view(1, 8, 3) << shift)
    key, inv = corners.reshape(-1, 3).unique(dim=0, return_inverse=True)
    return key.contiguous(), inv.reshape(-1, 8).contiguous()


class VoxelModel:
    """A loaded model: the tensors of the checkpoint on `device`, the derived voxel geometry and corner links,
    and `render`, the reference's `SVRenderer.render` for the per-voxel colour layout."""

    def __init__(self, state, device="cuda"):
        dev = torch.device(device)
        self.active_sh_degree = int(state["active_sh_degree"])
        self.scene_center = state["scene_center"].to(dev)
        self.inside_extent = state["inside_extent"].to(dev)
        self.scene_extent = state["scene_extent"].to(dev)
        self.octpath = state["octpath"].to(dev).reshape(-1, 1)
        self.octlevel = state["octlevel"].to(dev).to(torch.int8).reshape(-1, 1)
        self._geo_grid_pts = state["_geo_grid_pts"].to(dev)
        self._sh0 = state["_sh0"].to(dev)
        self._shs = state["_shs"].to(dev)
        self.vox_center, self.vox_size = voxel_geometry(self.octpath, self.octlevel, self.scene_center,
                                                        self.scene_extent)
        self.grid_pts_key, self.vox_key = grid_pts_link(self.octpath, self.octlevel)
        if self.grid_pts_key.shape[0] != self._geo_grid_pts.shape[0]:
            raise RuntimeError(f"checkpoint has {self._geo_grid_pts.shape[0]} grid points but its octree has "
                               f"{self.grid_pts_key.shape[0]}")

    @property
    def num_voxels(self):
        return self.octpath.shape[0]

    def vox_fn(self, idx, cam_pos, color_mode=None):
        """SRC/sparse_voxel_gears/renderer.py:59-153 for the per-voxel colour layout ('sh', 'shN')."""
        geos = renderer.GatherGeoParams.apply(self.vox_key, idx, self._geo_grid_pts)
        deg = self.active_sh_degree
        if color_mode is not None and color_mode != "sh":
            if not color_mode.startswith("sh"):
                raise NotImplementedError(color_mode)
            deg = int(color_mode[2])
        rgbs = renderer.SH_eval.apply(deg, idx, self.vox_center, cam_pos, None, self._sh0, self._shs)
        return {"geos": geos, "rgbs": rgbs,
                "subdiv_p": torch.ones([self.num_voxels, 1], dtype=torch.float32, device=self.octpath.device)}

    def render(self, camera, color_mode=None, n_samp_per_vox=1, bg_color=0.0, track_max_w=False,
               output_depth=False, output_normal=False, **other_opt):
        """One frame of `camera` (image_width, image_height, tanfovx, tanfovy, cx, cy, w2c, c2w, near):
        {'color', 'depth', 'normal', 'T', 'max_w'} (renderer.py:155-230 without super-sampling or
        background post-processing)."""
        settings = renderer.RasterSettings(
            color_mode=color_mode or "sh", n_samp_per_vox=n_samp_per_vox, image_width=camera.image_width,
            image_height=camera.image_height, tanfovx=camera.tanfovx, tanfovy=camera.tanfovy, cx=camera.cx,
            cy=camera.cy, w2c_matrix=camera.w2c, c2w_matrix=camera.c2w, bg_color=float(bg_color),
            near=getattr(camera, "near", 0.02), need_depth=output_depth, need_normal=output_normal,
            track_max_w=track_max_w, **other_opt)
        color, depth, normal, T, max_w = renderer.rasterize_voxels(settings, self.octpath, self.vox_center,
                                                                  self.vox_size, self.vox_fn)
        return {"color": color, "depth": depth if output_depth else None,
                "normal": normal if output_normal else None, "T": T, "max_w": max_w}


def load(path, device="cuda"):
    """A `VoxelModel` from a checkpoint file (io.py:49-96)."""
    return VoxelModel(read_state(path), device)


def save(path, model):
    """Writes `model` (a VoxelModel or a dict with the same tensor names) in the reference's unquantized
    format (io.py:17-46), loadable by the reference's `SVInOut.load`."""
    get = (lambda k: model[k]) if isinstance(model, dict) else (lambda k: getattr(model, k))
    sd = {"backend_name": FORMAT_BACKEND, "color_layout": "voxel", "geo_layout": "trilinear",
          "active_sh_degree": int(get("active_sh_degree"))}
    for k in ("scene_center", "inside_extent", "scene_extent", "octpath", "octlevel", "_geo_grid_pts", "_sh0",
              "_shs"):
        sd[k] = get(k).detach().contiguous().cpu()
    sd["quantized"] = False
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    torch.save(sd, path)
