"""ctypes binding of libu3d_hip.so (include/u3d_hip.h).

The library is built in-tree by ``univeral_3d_reconstruction_system_amd.build`` (or
``__graft_entry__.build()``). Importing this module never falls back to anything else: if the
library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libu3d_hip.so")


class Camera(C.Structure):
    """u3d_camera = RenderCamera (engine/include/u3d/splat/Rasterizer.h:13-21)."""
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("R", C.c_float * 9), ("t", C.c_float * 3)]


class RasterSettings(C.Structure):
    """u3d_raster_settings = RasterSettings (Rasterizer.h:23-32)."""
    _fields_ = [("tile_size", C.c_int), ("near_plane", C.c_float), ("eps2d", C.c_float),
                ("background", C.c_float * 3), ("min_alpha", C.c_float),
                ("min_transmittance", C.c_float)]


class Gaussians(C.Structure):
    """u3d_gaussians: device SoA pointers (GaussianModel.h:23-37)."""
    _fields_ = [("means", C.c_void_p), ("scales", C.c_void_p), ("quats", C.c_void_p),
                ("opacities", C.c_void_p), ("colors", C.c_void_p), ("n", C.c_int64)]


class LearningRates(C.Structure):
    _fields_ = [("means", C.c_float), ("scales", C.c_float), ("quats", C.c_float),
                ("opacities", C.c_float), ("colors", C.c_float)]


class LossValue(C.Structure):
    _fields_ = [("total", C.c_float), ("l1", C.c_float), ("ssim", C.c_float)]


class TrainOptions(C.Structure):
    """u3d_train_options = TrainOptions (Trainer.h:19-37)."""
    _fields_ = [("iterations", C.c_int), ("max_gaussians", C.c_int), ("ssim_lambda", C.c_float),
                ("lr", LearningRates), ("means_lr_final", C.c_float),
                ("densify_grad_threshold", C.c_float), ("percent_dense", C.c_float),
                ("refine_start", C.c_int), ("refine_stop", C.c_int), ("refine_interval", C.c_int),
                ("opacity_reset_interval", C.c_int), ("prune_opacity", C.c_float),
                ("seed", C.c_uint), ("raster", RasterSettings),
                # extensions: SH degree, SH rate scale, views per step, hipGraph capture
                ("sh_degree", C.c_int), ("sh_lr_scale", C.c_float), ("views_per_step", C.c_int),
                ("use_graph", C.c_int), ("batch_scaling", C.c_int)]


class BatchTuning(C.Structure):
    """u3d_batch_tuning: launch choices of the batched engine."""
    _fields_ = [("heavy_first", C.c_int), ("workgroups_per_cu", C.c_int), ("lds_batch", C.c_int),
                ("sort_items", C.c_int), ("tile_size", C.c_int), ("atomic_lanes", C.c_int),
                ("sort_radix", C.c_int), ("pair_sort", C.c_int)]


class TrainStats(C.Structure):
    _fields_ = [("iteration", C.c_int), ("loss", C.c_float), ("psnr", C.c_float),
                ("gaussians", C.c_int64)]


class RasterState(C.Structure):
    _fields_ = [("n", C.c_int64), ("num_pairs", C.c_int64), ("tiles_x", C.c_int),
                ("tiles_y", C.c_int), ("radius", C.c_void_p), ("rect", C.c_void_p),
                ("depth", C.c_void_p), ("mean2d", C.c_void_p), ("conic", C.c_void_p),
                ("color", C.c_void_p), ("opacity", C.c_void_p), ("flags", C.c_void_p),
                ("tile_offsets", C.c_void_p), ("tile_list", C.c_void_p), ("final_t", C.c_void_p),
                ("n_contrib", C.c_void_p)]


P = C.c_void_p
I = C.c_int
I64 = C.c_int64
F = C.c_float

# name -> (restype, argtypes)
SIGNATURES = {
    "u3d_last_error": (C.c_char_p, []),
    "u3d_version": (C.c_char_p, []),
    "u3d_ctx_create": (I, [I, C.POINTER(P)]),
    "u3d_ctx_destroy": (None, [P]),
    "u3d_ctx_synchronize": (I, [P]),
    "u3d_ctx_stream": (P, [P]),
    "u3d_malloc": (I, [P, C.c_size_t, C.POINTER(P)]),
    "u3d_free": (I, [P, P]),
    "u3d_memcpy_h2d": (I, [P, P, P, C.c_size_t]),
    "u3d_memcpy_d2h": (I, [P, P, P, C.c_size_t]),
    "u3d_memset": (I, [P, P, I, C.c_size_t]),
    "u3d_raster_create": (I, [P, C.POINTER(P)]),
    "u3d_raster_destroy": (None, [P]),
    "u3d_raster_forward": (I, [P, C.POINTER(Gaussians), C.POINTER(Camera), C.POINTER(RasterSettings), P]),
    "u3d_raster_backward": (I, [P, C.POINTER(Gaussians), P, C.POINTER(Gaussians), P]),
    "u3d_raster_forward_sh": (I, [P, C.POINTER(Gaussians), P, I, C.POINTER(Camera),
                                  C.POINTER(RasterSettings), P]),
    "u3d_raster_backward_sh": (I, [P, C.POINTER(Gaussians), P, I, P, C.POINTER(Gaussians), P, P]),
    "u3d_trainer_get_sh_rest": (I, [P, P]),
    "u3d_trainer_set_sh_rest": (I, [P, P]),
    "u3d_trainer_get_sh_moments": (I, [P, P, P]),
    "u3d_trainer_get_densify_stats": (I, [P, P, P]),
    "u3d_raster_radii": (I, [P, C.POINTER(P), C.POINTER(I64)]),
    "u3d_raster_get_state": (I, [P, C.POINTER(RasterState)]),
    "u3d_raster_get_grad2d": (I, [P, C.POINTER(P)]),
    "u3d_photometric_loss": (I, [P, P, P, I, I, F, P, C.POINTER(LossValue)]),
    "u3d_photometric_loss_noalloc": (I, [P, P, P, I, I, F, P, P]),
    "u3d_psnr": (I, [P, P, P, I64, C.POINTER(F)]),
    "u3d_adam_create": (I, [P, C.POINTER(P)]),
    "u3d_adam_destroy": (None, [P]),
    "u3d_adam_reset": (I, [P, I64]),
    "u3d_adam_compact": (I, [P, P]),
    "u3d_adam_grow": (I, [P, I64]),
    "u3d_adam_reset_opacity_moments": (I, [P]),
    "u3d_adam_step": (I, [P, C.POINTER(Gaussians), C.POINTER(Gaussians), C.POINTER(LearningRates)]),
    "u3d_adam_moments": (I, [P, C.POINTER(Gaussians), C.POINTER(Gaussians), C.POINTER(I)]),
    "u3d_densify_decide": (I, [P, C.POINTER(Gaussians), P, P, F, I, F, F, F, I, P, P,
                               C.POINTER(I64), C.POINTER(I64), C.POINTER(I64)]),
    "u3d_densify_apply": (I, [P, C.POINTER(Gaussians), P, P, P, F, C.POINTER(I64)]),
    "u3d_trainer_create": (I, [P, P, P, P, P, P, I64, P, P, I, F, C.POINTER(TrainOptions), C.POINTER(P)]),
    "u3d_trainer_destroy": (None, [P]),
    "u3d_trainer_step": (I, [P]),
    "u3d_trainer_done": (I, [P]),
    "u3d_trainer_iteration": (I, [P]),
    "u3d_trainer_size": (I64, [P]),
    "u3d_trainer_get_model": (I, [P, P, P, P, P, P]),
    "u3d_trainer_stats": (I, [P, C.POINTER(TrainStats)]),
    "u3d_trainer_evaluate_psnr": (I, [P, C.POINTER(F)]),
    "u3d_trainer_evaluate_l1": (I, [P, C.POINTER(F)]),
    "u3d_trainer_render": (I, [P, C.POINTER(Camera), P]),
    "u3d_ctx_set_timing": (I, [P, I]),
    "u3d_ctx_set_option": (I, [P, I, I]),
    "u3d_ctx_stage_times": (I, [P, C.POINTER(C.c_double), C.POINTER(I64)]),
    "u3d_trainer_counters": (I, [P, C.POINTER(I64)]),
    "u3d_trainer_get_moments": (I, [P, P, P]),
    "u3d_trainer_get_tuning": (I, [P, C.POINTER(BatchTuning)]),
    "u3d_trainer_set_tuning": (I, [P, C.POINTER(BatchTuning)]),
    "u3d_points_knn_init": (I, [P, P, P, I, I64, C.POINTER(Gaussians)]),
    "u3d_points_features": (I, [P, P, P, I, I64, P, I, F, P]),
    "u3d_init_net_create": (I, [P, C.c_uint, C.POINTER(P)]),
    "u3d_init_net_destroy": (None, [P]),
    "u3d_init_net_param_count": (I64, []),
    "u3d_init_net_get_params": (I, [P, P]),
    "u3d_init_net_set_params": (I, [P, P]),
    "u3d_init_net_get_grads": (I, [P, P]),
    "u3d_init_net_forward": (I, [P, P, I64, P]),
    "u3d_init_net_train_step": (I, [P, P, P, P, I64, F, C.POINTER(F)]),
    "u3d_points_learned_init": (I, [P, P, P, P, I, I64, P, I, F, C.POINTER(Gaussians)]),
    "u3d_write_splat_ply": (I, [C.c_char_p, P, P, P, P, P, I64, P, I]),
    "u3d_read_splat_ply": (I, [C.c_char_p, I64, P, P, P, P, P, P, C.POINTER(I64), C.POINTER(I)]),
    "u3d_ply_info": (I, [C.c_char_p, C.POINTER(I64), C.POINTER(I), C.POINTER(I)]),
    "u3d_write_points_ply": (I, [C.c_char_p, P, P, I64, I64]),
    "u3d_read_points_ply": (I, [C.c_char_p, I64, P, P, C.POINTER(I64)]),
    "u3d_trainer_export_ply": (I, [P, C.c_char_p]),
    "u3d_view_splat_attributes": (I, [P, C.POINTER(Gaussians), P, P]),
    "u3d_view_depth_order": (I, [P, P, I64, P, P, P, P, P]),
    "u3d_scene_set_create": (I, [P, I, C.c_size_t, C.POINTER(P)]),
    "u3d_scene_set_destroy": (None, [P]),
    "u3d_scene_set_add": (I, [P, P]),
    "u3d_scene_set_size": (I, [P]),
    "u3d_scene_set_member": (P, [P, I]),
    "u3d_scene_set_step": (I, [P]),
    "u3d_scene_set_done": (I, [P]),
    "u3d_scene_set_bytes": (C.c_size_t, [P]),
    "u3d_scene_set_set_tuning": (I, [P, C.POINTER(BatchTuning)]),
    "u3d_trainer_autotune": (I, [P, I, I, C.POINTER(BatchTuning)]),
    "u3d_scene_set_autotune": (I, [P, I, I, C.POINTER(BatchTuning)]),
    "u3d_expf_host_array": (None, [P, P, I64]),
}

STAGES = ("project", "sort", "raster_fwd", "loss", "raster_bwd", "chain_rule", "adam", "densify",
          "graph_step")
U3D_NUM_STAGES = len(STAGES)
OPT_SEGMENTED_BACKWARD = 1
OPT_SORT_SPIN_LIMIT = 2
FEATURE_DIM = 80
NET_OUT = 16
NET_LAYERS = ((80, 64), (64, 256), (256, 256), (256, 256), (256, 16))  # (in, out)
NET_GELU = (False, True, True, True, False)
COUNTERS = ("steps", "gaussians", "pixels", "pairs", "views", "overflow_steps", "evaluated_pairs")
U3D_NUM_COUNTERS = len(COUNTERS)

_lib = None


def load() -> C.CDLL:
    """Loads libu3d_hip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -m univeral_3d_reconstruction_system_amd.build`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class U3DError(RuntimeError):
    """A non-zero status of the C ABI; ``code`` is the U3D_ERR_* value."""

    def __init__(self, code: int, message: str):
        super().__init__(f"u3d error {code}: {message}")
        self.code = code


def check(rc: int) -> None:
    if rc != 0:
        msg = load().u3d_last_error().decode(errors="replace")
        raise U3DError(rc, msg)
