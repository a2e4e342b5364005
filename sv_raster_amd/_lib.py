"""ctypes binding of libsvraster_hip.so (declarations: include/svraster_hip.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SVR_LIB_PATH") or os.path.join(_HERE, "libsvraster_hip.so")

c_i32, c_i64, c_f32, c_f64, c_p, c_sz = (ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double,
                                          ctypes.c_void_p, ctypes.c_size_t)


class SvrConstants(ctypes.Structure):
    _fields_ = [("max_num_levels", ctypes.c_int32), ("step_sz_scale", ctypes.c_float),
                ("max_n_samp", ctypes.c_int32), ("max_sort_key_bits", ctypes.c_int32),
                ("max_packed_image_dim", ctypes.c_int32), ("max_render_tiles", ctypes.c_int64),
                ("tile_x", ctypes.c_int32), ("tile_y", ctypes.c_int32)]


class SvrView(ctypes.Structure):
    """svr_view of include/svraster_hip.h (152 bytes)."""
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("tanfovx", ctypes.c_float),
                ("tanfovy", ctypes.c_float), ("cx", ctypes.c_float), ("cy", ctypes.c_float),
                ("near_plane", ctypes.c_float), ("bg_color", ctypes.c_float), ("w2c", ctypes.c_float * 12),
                ("c2w", ctypes.c_float * 12), ("tile_base", ctypes.c_int32), ("tile_shape", ctypes.c_int32),
                ("bwd_window", ctypes.c_int32), ("bwd_sync", ctypes.c_int32), ("pix_base", ctypes.c_int64)]


assert ctypes.sizeof(SvrView) == 152

# name -> (restype, argtypes)
SIGNATURES = {
    "svr_last_error": (ctypes.c_char_p, []),
    "svr_get_constants": (c_i32, [ctypes.POINTER(SvrConstants)]),
    "svr_geom_state_size": (c_sz, [c_i64]),
    "svr_order_table_size": (c_sz, [c_i64]),
    "svr_order_table_scratch_size": (c_sz, [c_i64]),
    "svr_binning_state_size": (c_sz, [c_i64]),
    "svr_image_state_size": (c_sz, [c_i32, c_i32, c_i64]),
    "svr_rasterize_preprocess": (c_i32, [c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_p, c_p, c_f32, c_i64,
                                         c_p, c_p, c_p, c_p, c_p]),
    "svr_num_rendered": (c_i32, [c_i64, c_p, c_p, ctypes.POINTER(c_i64)]),
    "svr_visible_voxels": (c_i32, [c_i64, c_p, c_p, c_p, c_p, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    "svr_build_order_table": (c_i32, [c_i64, c_p, c_p, c_p, c_p]),
    "svr_rasterize_binning": (c_i32, [c_i32, c_i32, c_i64, c_p, c_p, c_i64, c_p, c_p, c_i32, c_p]),
    "svr_rasterize_forward": (c_i32, [c_i32, c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_p, c_f32, c_i64,
                                      c_p, c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_i32, c_p]),
    "svr_backward_workspace_size": (c_sz, [c_i64]),
    "svr_rasterize_backward": (c_i32, [c_i32, c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_p, c_f32, c_i64,
                                       c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                       c_p, c_p]),
    "svr_rasterize_backward_noalloc": (c_i32, [c_i32, c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_p, c_f32, c_i64,
                                               c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                               c_p, c_p]),
    "svr_rasterize_forward_ext": (c_i32, [c_i32, c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_p, c_f32, c_i32,
                                          c_i32, c_i32, c_i32, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_i64, c_p,
                                          c_p, c_p, c_p, c_p, c_p, c_p, c_i32, c_p]),
    "svr_rasterize_backward_ext": (c_i32, [c_i32, c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_p, c_f32, c_i64,
                                           c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                           c_f32, c_p, c_f32, c_f32, c_i32, c_i32, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_unpack_image_state": (c_i32, [c_i32, c_i32, c_p, c_p, c_p, c_p, c_p]),
    "svr_batch_rasterize_max_w": (c_i32, [c_i32, c_i32, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_p,
                                          c_p, c_p, c_p]),
    "svr_batch_training_stat": (c_i32, [c_i32, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_batch_mark_samp": (c_i32, [c_i32, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_unpack_geom_state": (c_i32, [c_i64, c_p, c_p, c_p, c_p, c_p]),
    "svr_export_sort_keys": (c_i32, [c_i32, c_i32, c_i64, c_p, c_p, c_p, c_i64, c_p, c_p, c_p, c_p]),
    "svr_sh_compute": (c_i32, [c_i32, c_i64, c_p, c_i64, c_p, c_p, c_p, c_p, c_i32, c_p, c_p]),
    "svr_sh_compute_bw": (c_i32, [c_i32, c_i32, c_i64, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_sh_compute_visible": (c_i32, [c_i32, c_i64, c_p, c_p, c_p, c_p, c_p, c_i32, c_p, c_p]),
    "svr_sh_compute_bw_visible": (c_i32, [c_i32, c_i32, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_unbiased_adam_step": (c_i32, [c_i32, c_i64, c_p, c_p, c_p, c_p, c_f64, c_f64, c_f64, c_f64, c_f32,
                                       c_p]),
    "svr_biased_adam_step": (c_i32, [c_i32, c_i64, c_p, c_p, c_p, c_p, c_f32, c_f32, c_f32, c_f32, c_p]),
    "svr_adam_step_multi": (c_i32, [c_i32, c_i32, c_i32, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_voxel_order_rank": (c_i32, [c_i64, c_p, c_p, c_p]),
    "svr_ijk_2_octpath": (c_i32, [c_i64, c_p, c_p, c_p, c_p]),
    "svr_octpath_2_ijk": (c_i32, [c_i64, c_p, c_p, c_p, c_p]),
    "svr_batch_geom_state_size": (c_sz, [c_i32, c_i64]),
    "svr_batch_image_state_size": (c_sz, [c_i64, c_i64, c_i64]),
    "svr_batch_backward_workspace_size": (c_sz, [c_i32, c_i64, c_i64]),
    "svr_batch_preprocess": (c_i32, [c_i32, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p]),
    "svr_batch_sh_compute": (c_i32, [c_i32, c_i32, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_batch_sh_compute_bw": (c_i32, [c_i32, c_i32, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_batch_rasterize_forward": (c_i32, [c_i32, c_i32, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_i64,
                                            c_p, c_p, c_p, c_p, c_p]),
    "svr_batch_rasterize_backward": (c_i32, [c_i32, c_i32, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_i64, c_p,
                                             c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_batch_unpack_image_state": (c_i32, [c_i64, c_i64, c_p, c_p, c_p, c_p, c_p]),
    "svr_scan_entries": (c_i32, [c_i64, c_p, c_p]),
    "svr_batch_rasterize_forward_capacity": (c_i32, [c_i32, c_i32, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p, c_p,
                                                     c_i64, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_batch_rasterize_backward_capacity": (c_i32, [c_i32, c_i32, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p,
                                                      c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "svr_unbiased_adam_step_device": (c_i32, [c_i32, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_f64, c_f64,
                                              c_f64, c_f32, c_p]),
    "svr_adam_bias_terms": (c_i32, [c_f64, c_f64, c_f64, c_f64, c_p, c_p]),
    "svr_gather_geo_params": (c_i32, [c_i64, c_p, c_p, c_p, c_p, c_p]),
    "svr_gather_geo_params_bw": (c_i32, [c_i64, c_p, c_p, c_p, c_p, c_p]),
    "svr_gather_feat_params": (c_i32, [c_i64, c_i32, c_p, c_p, c_p, c_p, c_p]),
    "svr_gather_feat_params_bw": (c_i32, [c_i64, c_i32, c_p, c_p, c_p, c_p, c_p]),
    "svr_total_variation_workspace_size": (c_sz, [c_i64, c_i32]),
    "svr_total_variation_bw": (c_i32, [c_i64, c_i32, c_p, c_p, c_f32, c_p, c_i32, c_i32, c_p, c_p, c_p]),
    "svr_is_in_cone": (c_i32, [c_i64, c_f32, c_f32, c_f32, c_p, c_p, c_p, c_p]),
    "svr_compute_rd": (c_i32, [c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_p, c_p, c_p]),
    "svr_depth2pts": (c_i32, [c_i32, c_i32, c_f32, c_f32, c_f32, c_f32, c_p, c_p, c_p, c_p]),
    "svr_filter_geom_state": (c_i32, [c_i64, c_i64, c_p, c_p, c_p, c_p]),
    "svr_debug_sync": (c_i32, [c_p]),
    "svr_profile_enable": (c_i32, [c_i32]),
    "svr_profile_collect": (c_i32, [c_i32, ctypes.POINTER(c_f64), ctypes.POINTER(c_i64)]),
}

_lib = None


def load():
    """Load the HIP library (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python sv_raster_amd/build.py` "
            "(hipcc --offload-arch=gfx950); sv_raster_amd has no CPU fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(SIGNATURES)


def check(status, what=""):
    if status != 0:
        msg = load().svr_last_error().decode(errors="replace")
        if status == 1:
            raise Exception(f"{what}: {msg}" if what else msg)
        if status == 3:
            raise NotImplementedError(f"{what}: {msg}")
        raise RuntimeError(f"{what} failed (status {status}): {msg}")
