"""ctypes bindings of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle_splat.so`` (the repository's restatement of the reference trainer,
engine/src/splat) and, when it has been built, ``oracle/_ref/libu3d_ref.so`` (the reference
itself, compiled from its own sources by oracle/Makefile). Only the test suite, the smoke
check and the CPU-baseline leg of bench.py import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# U3D_ORACLE_SO names another build of the same source (the sanitizer build of
# tests/sanitize/Makefile); the default is the plain build.
ORACLE_SO = os.environ.get("U3D_ORACLE_SO") or os.path.join(HERE, "liboracle_splat.so")
REF_SO = os.path.join(HERE, "_ref", "libu3d_ref.so")

F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class Camera(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("R", C.c_float * 9), ("t", C.c_float * 3)]


class Settings(C.Structure):
    _fields_ = [("tile_size", C.c_int), ("near_plane", C.c_float), ("eps2d", C.c_float),
                ("background", C.c_float * 3), ("min_alpha", C.c_float),
                ("min_transmittance", C.c_float)]


class Model(C.Structure):
    _fields_ = [("means", C.POINTER(C.c_float)), ("scales", C.POINTER(C.c_float)),
                ("quats", C.POINTER(C.c_float)), ("opacities", C.POINTER(C.c_float)),
                ("colors", C.POINTER(C.c_float)), ("n", C.c_int)]


class Lrs(C.Structure):
    _fields_ = [("means", C.c_float), ("scales", C.c_float), ("quats", C.c_float),
                ("opacities", C.c_float), ("colors", C.c_float)]


class TrainOptions(C.Structure):
    _fields_ = [("iterations", C.c_int), ("max_gaussians", C.c_int), ("ssim_lambda", C.c_float),
                ("lr", Lrs), ("means_lr_final", C.c_float), ("densify_grad_threshold", C.c_float),
                ("percent_dense", C.c_float), ("refine_start", C.c_int), ("refine_stop", C.c_int),
                ("refine_interval", C.c_int), ("opacity_reset_interval", C.c_int),
                ("prune_opacity", C.c_float), ("seed", C.c_uint), ("raster", Settings)]


class LossValue(C.Structure):
    _fields_ = [("total", C.c_float), ("l1", C.c_float), ("ssim", C.c_float)]


FIELDS = (("means", 3), ("scales", 3), ("quats", 4), ("opacities", 1), ("colors", 3))


@dataclass
class Gaussians:
    """Host SoA parameter set in the reference layout (GaussianModel.h:23-37)."""
    means: np.ndarray
    scales: np.ndarray
    quats: np.ndarray
    opacities: np.ndarray
    colors: np.ndarray

    @property
    def n(self) -> int:
        return int(self.opacities.shape[0])

    @staticmethod
    def zeros(n: int) -> "Gaussians":
        return Gaussians(*(np.zeros(n * d, np.float32) for _, d in FIELDS))

    def copy(self) -> "Gaussians":
        return Gaussians(*(getattr(self, f).copy() for f, _ in FIELDS))

    def arrays(self):
        return [getattr(self, f) for f, _ in FIELDS]

    def struct(self) -> Model:
        for f, _ in FIELDS:
            a = getattr(self, f)
            assert a.dtype == np.float32 and a.flags.c_contiguous, f
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        return Model(fp(self.means), fp(self.scales), fp(self.quats), fp(self.opacities),
                     fp(self.colors), self.n)


def make_camera(cam: dict) -> Camera:
    c = Camera()
    c.width, c.height = int(cam["width"]), int(cam["height"])
    c.fx, c.fy, c.cx, c.cy = (float(cam[k]) for k in ("fx", "fy", "cx", "cy"))
    c.R[:] = [float(v) for v in np.asarray(cam["R"], np.float32).reshape(9)]
    c.t[:] = [float(v) for v in np.asarray(cam["t"], np.float32).reshape(3)]
    return c


def make_settings(tile_size=16, near_plane=0.01, eps2d=0.3, background=(0.0, 0.0, 0.0),
                  min_alpha=1.0 / 255.0, min_transmittance=1e-4) -> Settings:
    s = Settings()
    s.tile_size = tile_size
    s.near_plane = np.float32(near_plane)
    s.eps2d = np.float32(eps2d)
    s.background[:] = [float(np.float32(b)) for b in background]
    s.min_alpha = np.float32(min_alpha)
    s.min_transmittance = np.float32(min_transmittance)
    return s


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_orc = None
_ref = None


def lib():
    global _orc
    if _orc is None:
        _orc = _load(ORACLE_SO)
        if _orc is None:
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle`)")
        L = _orc
        L.orc_expf.restype = C.c_float
        L.orc_expf.argtypes = [C.c_float]
        L.orc_expf_array.argtypes = [F32P, F32P, C.c_long]
        L.orc_raster_new.restype = C.c_void_p
        L.orc_raster_free.argtypes = [C.c_void_p]
        L.orc_raster_forward.argtypes = [C.c_void_p, C.POINTER(Model), C.POINTER(Camera),
                                         C.POINTER(Settings), F32P]
        L.orc_raster_backward.argtypes = [C.c_void_p, C.POINTER(Model), F32P, C.POINTER(Model),
                                          C.c_void_p]
        L.orc_raster_num_pairs.restype = C.c_long
        L.orc_raster_num_pairs.argtypes = [C.c_void_p]
        L.orc_raster_num_tiles.argtypes = [C.c_void_p]
        L.orc_raster_get_state.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.orc_raster_get_projection.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        L.orc_raster_get_grad2d.argtypes = [C.c_void_p, F32P]
        L.orc_project.argtypes = [C.POINTER(Model), C.POINTER(Camera), C.POINTER(Settings)] + [C.c_void_p] * 9
        L.orc_bin_tiles.restype = C.c_long
        L.orc_bin_tiles.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_long]
        D64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.orc_render_f64.argtypes = [C.c_int, D64, D64, D64, D64, D64, C.POINTER(Camera),
                                     C.POINTER(Settings), D64]
        L.orc_raster_forward_sh.argtypes = [C.c_void_p, C.POINTER(Model), C.c_void_p, C.c_int,
                                            C.POINTER(Camera), C.POINTER(Settings), F32P]
        L.orc_raster_backward_sh.argtypes = [C.c_void_p, C.POINTER(Model), C.c_void_p, C.c_int, F32P,
                                             C.POINTER(Model), C.c_void_p, C.c_void_p]
        L.orc_render_f64_sh.argtypes = [C.c_int, D64, D64, D64, D64, D64, D64, C.c_int,
                                        C.POINTER(Camera), C.POINTER(Settings), D64]
        L.orc_sh_basis.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, D64]
        L.orc_adam_array.argtypes = [F32P, F32P, F32P, F32P, C.c_long, C.c_float, C.c_int]
        L.orc_photometric_loss.restype = LossValue
        L.orc_photometric_loss.argtypes = [F32P, F32P, C.c_int, C.c_int, C.c_float, F32P]
        L.orc_psnr.restype = C.c_float
        L.orc_psnr.argtypes = [F32P, F32P, C.c_long]
        L.orc_adam_step.argtypes = [C.POINTER(Model)] * 4 + [C.POINTER(Lrs), C.c_int]
        L.orc_densify.restype = C.c_long
        L.orc_densify.argtypes = [C.POINTER(Model), F32P, C.c_void_p, C.c_float, C.c_int, C.c_float,
                                  C.c_float, C.c_float, C.c_int, C.c_void_p, C.c_long, C.c_void_p,
                                  C.c_void_p, C.POINTER(Model), C.c_long, C.POINTER(C.c_long)]
        L.orc_init_from_points.argtypes = [C.c_int, F32P, C.c_void_p, C.c_int, C.POINTER(Model)]
        L.orc_options_for_iterations.argtypes = [C.c_int, C.POINTER(TrainOptions)]
        L.orc_trainer_new.restype = C.c_void_p
        L.orc_trainer_new.argtypes = [C.POINTER(Model), C.POINTER(Camera), C.c_void_p, C.c_int,
                                      C.c_float, C.POINTER(TrainOptions)]
        for name in ("free", "step"):
            getattr(L, f"orc_trainer_{name}").argtypes = [C.c_void_p]
        for name in ("done", "iteration", "size"):
            getattr(L, f"orc_trainer_{name}").argtypes = [C.c_void_p]
            getattr(L, f"orc_trainer_{name}").restype = C.c_int
        L.orc_trainer_get_model.argtypes = [C.c_void_p, C.POINTER(Model)]
        L.orc_trainer_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.POINTER(C.c_long)]
        for name in ("evaluate_psnr", "evaluate_l1"):
            getattr(L, f"orc_trainer_{name}").argtypes = [C.c_void_p]
            getattr(L, f"orc_trainer_{name}").restype = C.c_float
        L.orc_trainer_render.argtypes = [C.c_void_p, C.POINTER(Camera), F32P]
        L.orc_project_f64.argtypes = [C.c_int] + [C.c_void_p] * 6 + [C.c_int, C.POINTER(Camera),
                                                                   C.POINTER(Settings)] + [C.c_void_p] * 5
        L.orc_chain_rule_f64.argtypes = [C.c_void_p, C.POINTER(Model), C.c_void_p, C.c_int, F32P] + \
            [C.c_void_p] * 6
        L.orc_view_attributes.argtypes = [C.POINTER(Model), C.c_void_p, C.c_void_p]
        L.orc_view_depth_order.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]
        L.orc_trainer_set_sh.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_void_p]
        L.orc_trainer_get_sh.argtypes = [C.c_void_p, F32P]
    return _orc


def ref_lib():
    """The compiled reference (None where oracle/_ref was not built)."""
    global _ref
    if _ref is None:
        _ref = _load(REF_SO)
        if _ref is None:
            return None
        L = _ref
        L.ref_raster_new.restype = C.c_void_p
        L.ref_raster_new.argtypes = [C.c_int]
        L.ref_raster_free.argtypes = [C.c_void_p]
        L.ref_raster_forward.argtypes = [C.c_void_p, C.POINTER(Model), C.POINTER(Camera),
                                         C.POINTER(Settings), F32P]
        L.ref_raster_num_pairs.restype = C.c_long
        L.ref_raster_num_pairs.argtypes = [C.c_void_p]
        L.ref_raster_get_state.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.ref_raster_get_projection.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        L.ref_raster_backward.argtypes = [C.c_void_p, C.POINTER(Model), F32P, C.POINTER(Model),
                                          C.c_void_p]
        L.ref_photometric_loss.restype = LossValue
        L.ref_photometric_loss.argtypes = [F32P, F32P, C.c_int, C.c_int, C.c_float, F32P, C.c_int]
        L.ref_psnr.restype = C.c_float
        L.ref_psnr.argtypes = [F32P, F32P, C.c_long]
        L.ref_adam_steps.argtypes = [C.POINTER(Model), C.POINTER(Model), C.POINTER(Lrs), C.c_int,
                                     C.c_void_p, C.c_void_p]
        L.ref_init_from_points.argtypes = [C.c_int, F32P, C.c_void_p, C.c_int, C.POINTER(Model)]
        L.ref_write_splat_ply.restype = C.c_int
        L.ref_write_splat_ply.argtypes = [C.c_char_p, C.POINTER(Model)]
        L.ref_read_splat_ply.restype = C.c_long
        L.ref_read_splat_ply.argtypes = [C.c_char_p, C.c_void_p]
        L.ref_write_points_ply.restype = C.c_int
        L.ref_write_points_ply.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long]
        L.ref_read_points_ply.restype = C.c_long
        L.ref_read_points_ply.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p]
        L.ref_read_splat_ply_any.restype = C.c_long
        L.ref_read_splat_ply_any.argtypes = [C.c_char_p, C.c_void_p, C.POINTER(C.c_int)]
        L.ref_trainer_new.restype = C.c_void_p
        L.ref_trainer_new.argtypes = [C.POINTER(Model), C.POINTER(Camera), C.c_void_p, C.c_int,
                                      C.c_float, C.POINTER(TrainOptions), C.c_int]
        for name in ("free", "step"):
            getattr(L, f"ref_trainer_{name}").argtypes = [C.c_void_p]
        for name in ("done", "size"):
            getattr(L, f"ref_trainer_{name}").argtypes = [C.c_void_p]
            getattr(L, f"ref_trainer_{name}").restype = C.c_int
        L.ref_trainer_get_model.argtypes = [C.c_void_p, C.POINTER(Model)]
        L.ref_trainer_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.POINTER(C.c_long)]
        L.ref_trainer_evaluate_psnr.argtypes = [C.c_void_p]
        L.ref_trainer_evaluate_psnr.restype = C.c_float
        L.ref_trainer_densify.argtypes = [C.c_void_p, F32P, C.c_void_p, C.c_int]
        L.ref_normals.argtypes = [C.c_uint, C.c_int, F32P]
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------------------- rasterizer
class Raster:
    """The oracle's Rasterizer (Rasterizer.h:38-90); ``use_ref`` selects the compiled reference."""

    def __init__(self, use_ref: bool = False, threads: int = 1):
        self.use_ref = use_ref
        if use_ref:
            self.L = ref_lib()
            if self.L is None:
                raise RuntimeError("compiled reference not available")
            self.h = self.L.ref_raster_new(threads)
        else:
            self.L = lib()
            self.h = self.L.orc_raster_new()
        self.n = 0
        self.cam = None

    def __del__(self):
        try:
            if self.use_ref:
                self.L.ref_raster_free(self.h)
            else:
                self.L.orc_raster_free(self.h)
        except Exception:
            pass

    def forward(self, g: Gaussians, cam: dict, st: Settings, sh_rest=None, sh_degree=0) -> np.ndarray:
        c = make_camera(cam)
        rgb = np.zeros(c.width * c.height * 3, np.float32)
        m = g.struct()
        if sh_rest is not None:
            assert not self.use_ref, "the reference has no SH beyond degree 0"
            self._rest = np.ascontiguousarray(sh_rest, np.float32)
            self.L.orc_raster_forward_sh(self.h, C.byref(m), _ptr(self._rest), sh_degree, C.byref(c),
                                         C.byref(st), rgb)
        else:
            fn = self.L.ref_raster_forward if self.use_ref else self.L.orc_raster_forward
            fn(self.h, C.byref(m), C.byref(c), C.byref(st), rgb)
        self.n, self.cam, self.st = g.n, c, st
        return rgb

    def backward_sh(self, g: Gaussians, dl: np.ndarray, sh_rest, sh_degree, grads=None, ndc=None):
        """Backward of an SH forward; returns (grads, grad_sh_rest)."""
        if grads is None:
            grads = Gaussians.zeros(g.n)
        rest = np.ascontiguousarray(sh_rest, np.float32)
        grest = np.zeros(g.n * 45, np.float32)
        m, gm = g.struct(), grads.struct()
        self.L.orc_raster_backward_sh(self.h, C.byref(m), _ptr(rest), sh_degree,
                                      np.ascontiguousarray(dl, np.float32), C.byref(gm), _ptr(grest),
                                      _ptr(ndc))
        return grads, grest

    def num_pairs(self) -> int:
        return int((self.L.ref_raster_num_pairs if self.use_ref else self.L.orc_raster_num_pairs)(self.h))

    def state(self):
        ts = self.st.tile_size
        tiles = ((self.cam.width + ts - 1) // ts) * ((self.cam.height + ts - 1) // ts)
        P = self.cam.width * self.cam.height
        radius = np.zeros(self.n, np.int32)
        offsets = np.zeros(tiles + 1, np.int32)
        tl = np.zeros(self.num_pairs(), np.int32)
        finalT = np.zeros(P, np.float32)
        ncon = np.zeros(P, np.int32)
        fn = self.L.ref_raster_get_state if self.use_ref else self.L.orc_raster_get_state
        fn(self.h, _ptr(radius), _ptr(offsets), _ptr(tl), _ptr(finalT), _ptr(ncon))
        return dict(radius=radius, tile_offsets=offsets, tile_list=tl, final_t=finalT, n_contrib=ncon)

    def projection(self):
        n = self.n
        out = dict(mean2d=np.zeros(2 * n, np.float32), conic=np.zeros(3 * n, np.float32),
                   depth=np.zeros(n, np.float32), rgb=np.zeros(3 * n, np.float32),
                   alpha=np.zeros(n, np.float32), rect=np.zeros(4 * n, np.int32),
                   flags=np.zeros(n, np.uint8))
        fn = self.L.ref_raster_get_projection if self.use_ref else self.L.orc_raster_get_projection
        fn(self.h, *(_ptr(out[k]) for k in ("mean2d", "conic", "depth", "rgb", "alpha", "rect", "flags")))
        return out

    def backward(self, g: Gaussians, dl: np.ndarray, grads: Gaussians | None = None, ndc=None):
        if grads is None:
            grads = Gaussians.zeros(g.n)
        m, gm = g.struct(), grads.struct()
        fn = self.L.ref_raster_backward if self.use_ref else self.L.orc_raster_backward
        fn(self.h, C.byref(m), np.ascontiguousarray(dl, np.float32), C.byref(gm), _ptr(ndc))
        return grads

    def grad2d(self):
        assert not self.use_ref
        out = np.zeros(self.n * 9, np.float32)
        self.L.orc_raster_get_grad2d(self.h, out)
        return out

    def chain_rule_f64(self, g: "Gaussians", grad2d, sh_rest=None, sh_degree=0) -> dict:
        """The chain-rule stage of the backward in FP64 from 2D gradients [9n] (e.g. grad2d()
        after a backward) and this rasterizer's last forward; fields of the model plus 'sh'."""
        assert not self.use_ref
        n = g.n
        out = {f: np.zeros(n * d, np.float64) for f, d in FIELDS}
        out["sh"] = np.zeros(n * 45, np.float64)
        rest = None if sh_rest is None else np.ascontiguousarray(sh_rest, np.float32)
        m = g.struct()
        self.L.orc_chain_rule_f64(self.h, C.byref(m), _ptr(rest), sh_degree,
                                  np.ascontiguousarray(grad2d, np.float32),
                                  *(_ptr(out[f]) for f, _ in FIELDS), _ptr(out["sh"]))
        return out


def render_f64(g: dict, cam: dict, st: Settings) -> np.ndarray:
    c = make_camera(cam)
    rgb = np.zeros(c.width * c.height * 3, np.float64)
    arr = [np.ascontiguousarray(g[f], np.float64) for f, _ in FIELDS]
    lib().orc_render_f64(len(arr[3]), *arr, C.byref(c), C.byref(st), rgb)
    return rgb


def render_f64_sh(g: dict, rest, degree, cam: dict, st: Settings) -> np.ndarray:
    c = make_camera(cam)
    rgb = np.zeros(c.width * c.height * 3, np.float64)
    arr = [np.ascontiguousarray(g[f], np.float64) for f, _ in FIELDS]
    lib().orc_render_f64_sh(len(arr[3]), *arr, np.ascontiguousarray(rest, np.float64), degree,
                            C.byref(c), C.byref(st), rgb)
    return rgb


def project_f64(g: dict, cam: dict, st: Settings, sh_rest=None, sh_degree=0) -> dict:
    """The projection in FP64 (double parameters): mean2d, conic, rgb, alpha, radius."""
    n = len(g["opacities"])
    arr = [np.ascontiguousarray(g[f], np.float64) for f, _ in FIELDS]
    rest = None if sh_rest is None else np.ascontiguousarray(sh_rest, np.float64)
    out = dict(mean2d=np.zeros(2 * n), conic=np.zeros(3 * n), rgb=np.zeros(3 * n), alpha=np.zeros(n),
               radius=np.zeros(n, np.int32))
    c = make_camera(cam)
    lib().orc_project_f64(n, *(_ptr(a) for a in arr), _ptr(rest), sh_degree, C.byref(c), C.byref(st),
                          *(_ptr(out[k]) for k in ("mean2d", "conic", "rgb", "alpha", "radius")))
    return out


def view_attributes(g: Gaussians):
    """SplatView.cpp:117-180: (rgba [4n] uint8, cov6 [6n] float32)."""
    rgba = np.zeros(4 * g.n, np.uint8)
    cov = np.zeros(6 * g.n, np.float32)
    lib().orc_view_attributes(C.byref(g.struct()), _ptr(rgba), _ptr(cov))
    return rgba, cov


def view_depth_order(centers, row2) -> np.ndarray:
    """SplatView.cpp:349-397: front-to-back order of the splats for a view."""
    c = np.ascontiguousarray(centers, np.float32)
    r = np.ascontiguousarray(row2, np.float32)
    order = np.zeros(len(c) // 3, np.uint32)
    lib().orc_view_depth_order(_ptr(c), len(c) // 3, _ptr(r), _ptr(order))
    return order


def sh_basis(x, y, z, degree=3) -> np.ndarray:
    out = np.zeros(15, np.float64)
    lib().orc_sh_basis(x, y, z, degree, out)
    return out


def project(g: Gaussians, cam: dict, st: Settings):
    n = g.n
    out = dict(mean2d=np.zeros(2 * n, np.float32), conic=np.zeros(3 * n, np.float32),
               cov2d=np.zeros(3 * n, np.float32), depth=np.zeros(n, np.float32),
               rgb=np.zeros(3 * n, np.float32), alpha=np.zeros(n, np.float32),
               radius=np.zeros(n, np.int32), rect=np.zeros(4 * n, np.int32), flags=np.zeros(n, np.uint8))
    c, m = make_camera(cam), g.struct()
    lib().orc_project(C.byref(m), C.byref(c), C.byref(st),
                      *(_ptr(out[k]) for k in ("mean2d", "conic", "cov2d", "depth", "rgb", "alpha",
                                               "radius", "rect", "flags")))
    return out


def bin_tiles(radius, rect, depth, tiles_x, tiles_y):
    n = len(radius)
    cap = int(np.sum(np.where(radius > 0, (rect[2::4] - rect[0::4]) * (rect[3::4] - rect[1::4]), 0)))
    offsets = np.zeros(tiles_x * tiles_y + 1, np.int32)
    tl = np.zeros(max(cap, 1), np.int32)
    k = lib().orc_bin_tiles(n, _ptr(np.ascontiguousarray(radius, np.int32)),
                            _ptr(np.ascontiguousarray(rect, np.int32)),
                            _ptr(np.ascontiguousarray(depth, np.float32)), tiles_x, tiles_y,
                            _ptr(offsets), _ptr(tl), len(tl))
    return offsets, tl[:k]


# ------------------------------------------------------------------------------- loss
def photometric_loss(render, target, W, H, lam=0.2, use_ref=False, threads=1):
    render = np.ascontiguousarray(render, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    grad = np.zeros_like(render)
    if use_ref:
        v = ref_lib().ref_photometric_loss(render, target, W, H, lam, grad, threads)
    else:
        v = lib().orc_photometric_loss(render, target, W, H, lam, grad)
    return (v.total, v.l1, v.ssim), grad


def psnr(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return lib().orc_psnr(a, b, a.size)


# ------------------------------------------------------------------------------- adam
def lrs(means=1.6e-4, scales=5e-3, quats=1e-3, opacities=5e-2, colors=2.5e-3) -> Lrs:
    return Lrs(means, scales, quats, opacities, colors)


def adam_steps(p: Gaussians, g: Gaussians, lr: Lrs, steps: int, use_ref=False):
    """`steps` Adam steps with fixed gradients from zero moments; returns (p, m, v)."""
    p = p.copy()
    m, v = Gaussians.zeros(p.n), Gaussians.zeros(p.n)
    if use_ref:
        pm, gm, mm, vm = p.struct(), g.struct(), m.struct(), v.struct()
        ref_lib().ref_adam_steps(C.byref(pm), C.byref(gm), C.byref(lr), steps, C.byref(mm), C.byref(vm))
    else:
        pm, gm, mm, vm = p.struct(), g.struct(), m.struct(), v.struct()
        for s in range(1, steps + 1):
            lib().orc_adam_step(C.byref(pm), C.byref(gm), C.byref(mm), C.byref(vm), C.byref(lr), s)
    return p, m, v


# --------------------------------------------------------------------------- densify
def densify(g: Gaussians, accum, count, extent, max_gaussians, normals, grad_threshold=2e-4,
            percent_dense=0.01, prune_opacity=0.005, opacity_was_reset=False):
    n = g.n
    keep = np.zeros(n, np.uint8)
    kind = np.zeros(n, np.uint8)
    cap = 2 * n + 2
    extra = Gaussians.zeros(cap)
    used = C.c_long(0)
    normals = np.ascontiguousarray(normals, np.float32)
    m, em = g.struct(), extra.struct()
    cnt = lib().orc_densify(C.byref(m), np.ascontiguousarray(accum, np.float32),
                            _ptr(np.ascontiguousarray(count, np.int32)), extent, max_gaussians,
                            grad_threshold, percent_dense, prune_opacity, int(opacity_was_reset),
                            _ptr(normals), len(normals), _ptr(keep), _ptr(kind), C.byref(em), cap,
                            C.byref(used))
    extra = Gaussians(extra.means[:3 * cnt].copy(), extra.scales[:3 * cnt].copy(),
                      extra.quats[:4 * cnt].copy(), extra.opacities[:cnt].copy(),
                      extra.colors[:3 * cnt].copy())
    return keep, kind, extra, int(used.value)


def init_from_points(xyz, rgb=None, use_ref=False) -> Gaussians:
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    n = len(xyz) // 3
    out = Gaussians.zeros(n)
    om = out.struct()
    rgbp = None if rgb is None else np.ascontiguousarray(rgb, np.float32).reshape(-1)
    cnt = 0 if rgbp is None else len(rgbp)
    if use_ref:
        ref_lib().ref_init_from_points(n, xyz, _ptr(rgbp), cnt, C.byref(om))
    else:
        lib().orc_init_from_points(n, xyz, _ptr(rgbp), cnt, C.byref(om))
    return out


def options_for_iterations(iterations: int) -> TrainOptions:
    o = TrainOptions()
    lib().orc_options_for_iterations(iterations, C.byref(o))
    return o


# ---------------------------------------------------------------------------- trainer
class Trainer:
    """Oracle Trainer (Trainer.h:249-292); ``use_ref`` drives the compiled reference instead."""

    def __init__(self, init: Gaussians, cams: list, images_u8: list, extent: float,
                 opt: TrainOptions, use_ref=False, threads=1, sh_degree=0, sh_lr_scale=0.0,
                 sh_rest=None):
        self.use_ref = use_ref
        self.L = ref_lib() if use_ref else lib()
        arr = (Camera * len(cams))(*[make_camera(c) for c in cams])
        self._rgb = np.ascontiguousarray(np.concatenate([np.asarray(i, np.uint8).reshape(-1) for i in images_u8]))
        m = init.struct()
        if use_ref:
            self.h = self.L.ref_trainer_new(C.byref(m), arr, _ptr(self._rgb), len(cams), extent, C.byref(opt), threads)
        else:
            self.h = self.L.orc_trainer_new(C.byref(m), arr, _ptr(self._rgb), len(cams), extent, C.byref(opt))
        self.p = "ref_trainer_" if use_ref else "orc_trainer_"
        self.sh_degree = sh_degree
        if sh_degree:
            assert not use_ref, "the reference has no SH beyond degree 0"
            self._rest0 = None if sh_rest is None else np.ascontiguousarray(sh_rest, np.float32)
            self.L.orc_trainer_set_sh(self.h, sh_degree, sh_lr_scale, _ptr(self._rest0))

    def sh_rest(self) -> np.ndarray:
        out = np.zeros(self.size() * 45, np.float32)
        self.L.orc_trainer_get_sh(self.h, out)
        return out

    def _f(self, name):
        return getattr(self.L, self.p + name)

    def __del__(self):
        try:
            self._f("free")(self.h)
        except Exception:
            pass

    def step(self):
        self._f("step")(self.h)

    def done(self) -> bool:
        return bool(self._f("done")(self.h))

    def size(self) -> int:
        return int(self._f("size")(self.h))

    def model(self) -> Gaussians:
        out = Gaussians.zeros(self.size())
        om = out.struct()
        self._f("get_model")(self.h, C.byref(om))
        return out

    def stats(self):
        it, loss, ps, ng = C.c_int(), C.c_float(), C.c_float(), C.c_long()
        self._f("stats")(self.h, C.byref(it), C.byref(loss), C.byref(ps), C.byref(ng))
        return dict(iteration=it.value, loss=loss.value, psnr=ps.value, gaussians=ng.value)

    def evaluate_psnr(self) -> float:
        return float(self._f("evaluate_psnr")(self.h))

    def evaluate_l1(self) -> float:
        assert not self.use_ref
        return float(self.L.orc_trainer_evaluate_l1(self.h))

    def densify(self, accum, count, opacity_was_reset=False):
        assert self.use_ref
        self.L.ref_trainer_densify(self.h, np.ascontiguousarray(accum, np.float32),
                                   _ptr(np.ascontiguousarray(count, np.int32)), int(opacity_was_reset))
