"""The C-ABI library loads without a GPU, exports every entry point declared
in include/svraster_hip.h, and the host-side validation of the drop-in API
behaves like the reference's (no device work is launched by these tests)."""
import ctypes
import os
import re

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "svraster_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(svr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from sv_raster_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.SIGNATURES) == declared


def test_constants_match_reference_config():
    from sv_raster_amd import meta
    assert meta.MAX_NUM_LEVELS == 21
    assert meta.STEP_SZ_SCALE == 100.0
    assert meta.MAX_N_SAMP == 3
    assert meta.MAX_SORT_KEY_BITS == 128
    assert meta.MAX_PACKED_IMAGE_DIM == 65535
    assert meta.MAX_RENDER_TILES == ((65535 + 16) // 16) ** 2


def test_state_size_queries():
    from sv_raster_amd import _lib
    L = _lib.load()
    assert L.svr_geom_state_size(0) == 256 or L.svr_geom_state_size(0) >= 0
    assert L.svr_geom_state_size(1_000_000) >= 1_000_000 * (4 + 4 + 8 + 4 + 4)
    assert L.svr_image_state_size(1920, 1080, 0) >= 1920 * 1080 * 4 + 8160 * 12
    assert L.svr_image_state_size(1920, 1080, 64000) >= L.svr_image_state_size(1920, 1080, 0) + (1000 + 8160) * 2048
    assert L.svr_binning_state_size(1000) >= 1000 * (8 + 8 + 4 + 4 + 4)
    assert L.svr_order_table_size(1000) >= 8000 * 8
    assert L.svr_backward_workspace_size(1000) >= 1000 * 48


def test_invalid_arguments_fail_without_device_work():
    from sv_raster_amd import _lib
    L = _lib.load()
    rc = L.svr_rasterize_preprocess(0, 10, 0.5, 0.5, 5.0, 5.0, None, None, 0.02, 10, None, None, None, None, None)
    assert rc == 1 and b"image size" in L.svr_last_error()
    rc = L.svr_rasterize_forward(4, 16, 16, 0.5, 0.5, 8.0, 8.0, None, 0.0, 0, None, None, None, None, None, None, 0,
                                 None, None, None, None, 0, None)
    assert rc == 1 and b"n_samp_per_vox" in L.svr_last_error()
    with pytest.raises(Exception, match="image size"):
        _lib.check(1 if L.svr_rasterize_preprocess(70000, 10, 0.5, 0.5, 5.0, 5.0, None, None, 0.02, 1, None, None,
                                                   None, None, None) else 0, "rasterize_preprocess")


def _settings(**kw):
    from sv_raster_amd import renderer
    base = dict(color_mode="sh", n_samp_per_vox=1, image_width=16, image_height=16, tanfovx=0.5, tanfovy=0.5,
                cx=8.0, cy=8.0, w2c_matrix=torch.eye(4), c2w_matrix=torch.eye(4))
    base.update(kw)
    return renderer.RasterSettings(**base)


def test_rasterize_voxels_validation_matches_reference():
    from sv_raster_amd import renderer
    paths = torch.zeros((4, 1), dtype=torch.int64)
    centers = torch.zeros((4, 3))
    lengths = torch.ones((4, 1))
    with pytest.raises(Exception, match="Expect RasterSettings as first argument."):
        renderer.rasterize_voxels((), paths, centers, lengths, None)
    with pytest.raises(Exception, match=r"n_samp_per_vox should be in range \[1, 3\]."):
        renderer.rasterize_voxels(_settings(n_samp_per_vox=4), paths, centers, lengths, None)
    with pytest.raises(Exception, match="Size mismatched."):
        renderer.rasterize_voxels(_settings(), paths, centers[:3], lengths, None)
    with pytest.raises(Exception, match=r"Expect vox_centers in shape \[N, 3\]."):
        renderer.rasterize_voxels(_settings(), paths, torch.zeros((4, 2)), lengths, None)


def test_cpu_tensors_are_rejected_not_computed():
    """There is no CPU fallback: the preprocess refuses host tensors."""
    from sv_raster_amd import _C
    with pytest.raises(RuntimeError, match="HIP device"):
        _C.rasterize_preprocess(16, 16, 0.5, 0.5, 8.0, 8.0, torch.eye(4), torch.eye(4), 0.02,
                                torch.zeros((4, 1), dtype=torch.int64), torch.zeros((4, 3)), torch.ones((4, 1)))


def test_sparse_adam_constructor_checks():
    from sv_raster_amd.sparse_adam import SparseAdam
    p = torch.zeros(4, requires_grad=True)
    with pytest.raises(ValueError, match="Invalid learning rate"):
        SparseAdam([p], lr=-1.0)
    with pytest.raises(ValueError, match="Invalid beta parameter at index 0"):
        SparseAdam([p], betas=(1.0, 0.9))
    opt = SparseAdam([p], lr=0.1, betas=(0.1, 0.99), eps=1e-15, sparse=True)
    assert opt.sparse and not opt.biased


def test_ext_backward_checks_shapes_before_launch():
    """The optional-term backward rejects missing or mis-shaped inputs on the host, before any launch."""
    from sv_raster_amd import _C
    z, e = torch.zeros, torch.empty(0)
    args = dict(R=10, n_samp_per_vox=1, image_width=16, image_height=16, tan_fovx=0.5, tan_fovy=0.5, cx=8.0,
                cy=8.0, w2c_matrix=torch.eye(4), c2w_matrix=torch.eye(4), bg_color=0.0,
                octree_paths=z((4, 1), dtype=torch.int64), vox_centers=z((4, 3)), vox_lengths=torch.ones((4, 1)),
                geos=z((4, 8)), rgbs=z((4, 3)), geomBuffer=e, binningBuffer=e, imageBuffer=e,
                out_T=torch.ones((1, 16, 16)), dL_dout_color=z((3, 16, 16)), dL_dout_depth=e, dL_dout_normal=e,
                dL_dout_T=z((1, 16, 16)), lambda_R_concen=0.0, gt_color=e, lambda_ascending=0.0, lambda_dist=0.0,
                need_depth=False, need_normal=False, out_D=e, out_N=e)
    cases = [(dict(need_depth=True, dL_dout_depth=z((3, 8, 8))), "dL_dout_depth"),
             (dict(need_normal=True), "dL_dout_normal"),
             (dict(lambda_R_concen=0.1), "gt_color"),
             (dict(lambda_dist=0.1), "out_depth")]
    for over, what in cases:
        with pytest.raises(RuntimeError, match=what):
            _C.rasterize_voxels_backward(**dict(args, **over))


def test_batch_rejects_optional_outputs():
    from sv_raster_amd import batch, renderer
    s = renderer.RasterSettings(color_mode="sh", n_samp_per_vox=1, image_width=16, image_height=16, tanfovx=0.5,
                                tanfovy=0.5, cx=8.0, cy=8.0, w2c_matrix=torch.eye(4), c2w_matrix=torch.eye(4),
                                need_depth=True)
    with pytest.raises(NotImplementedError, match="need_depth"):
        batch.ViewBatch([s])
