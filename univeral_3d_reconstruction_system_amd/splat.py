"""Python mirror of the reference trainer API (engine/include/u3d/splat/*.h) over libu3d_hip.

Names and argument meaning follow the reference classes: ``Rasterizer.forward/backward/radii``
(Rasterizer.h:38-90), ``photometric_loss`` (Loss.h:18-19), ``Adam`` (Adam.h:18-33),
``Trainer`` (Trainer.h:249-292). Arrays cross the boundary as device buffers owned by
``DeviceArray``; host data is numpy. Every call goes through the C ABI and raises
``U3DError`` on failure -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _abi
from ._abi import U3DError, check

FIELDS = (("means", 3), ("scales", 3), ("quats", 4), ("opacities", 1), ("colors", 3))


class Context:
    """u3d_ctx: one HIP device and stream. Closing it first releases every library object
    created on it that is still alive, so no C object outlives its context."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load()
        self._children = weakref.WeakSet()
        h = C.c_void_p()
        check(self.lib.u3d_ctx_create(device, C.byref(h)))
        self.h = h

    def synchronize(self):
        check(self.lib.u3d_ctx_synchronize(self.h))

    def set_option(self, option: int, value: int):
        check(self.lib.u3d_ctx_set_option(self.h, int(option), int(value)))

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        if getattr(self, "h", None):
            # scene sets first: they destroy their members
            kids = sorted(list(self._children), key=lambda o: not isinstance(o, SceneSet))
            for obj in kids:
                obj._release()
            self.lib.u3d_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation of a fixed dtype and shape, filled from / copied to numpy."""

    def __init__(self, ctx: Context, shape, dtype=np.float32):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(ctx.lib.u3d_malloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value
        ctx._adopt(self)

    @staticmethod
    def from_numpy(ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(ctx, a.shape, a.dtype)
        if a.nbytes:
            check(ctx.lib.u3d_memcpy_h2d(ctx.h, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(self.ctx.lib.u3d_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self):
        check(self.ctx.lib.u3d_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def _release(self):
        if getattr(self, "ptr", None) and self.ctx.h:
            self.ctx.lib.u3d_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


def _view_numpy(ctx: Context, ptr: int, count: int, dtype) -> np.ndarray:
    dtype = np.dtype(dtype)
    out = np.empty(count, dtype)
    if count:
        check(ctx.lib.u3d_memcpy_d2h(ctx.h, out.ctypes.data, ptr, out.nbytes))
    return out


class DeviceGaussians:
    """GaussianModel on the device (also used for gradients): five SoA arrays."""

    def __init__(self, ctx: Context, n: int, capacity: int | None = None):
        self.ctx = ctx
        self.n = int(n)
        cap = max(int(capacity or n), 1)
        self.arrays = {f: DeviceArray(ctx, (cap * d,)) for f, d in FIELDS}
        for a in self.arrays.values():
            a.zero()

    @staticmethod
    def from_host(ctx: Context, g, capacity: int | None = None) -> "DeviceGaussians":
        get = (lambda f: g[f]) if isinstance(g, dict) else (lambda f: getattr(g, f))
        n = len(np.asarray(get("opacities")))
        d = DeviceGaussians(ctx, n, capacity)
        for f, _ in FIELDS:
            a = np.ascontiguousarray(get(f), np.float32).reshape(-1)
            if a.nbytes:
                check(ctx.lib.u3d_memcpy_h2d(ctx.h, d.arrays[f].ptr, a.ctypes.data, a.nbytes))
        return d

    def struct(self) -> _abi.Gaussians:
        a = self.arrays
        return _abi.Gaussians(a["means"].ptr, a["scales"].ptr, a["quats"].ptr,
                              a["opacities"].ptr, a["colors"].ptr, self.n)

    def to_host(self) -> dict:
        return {f: _view_numpy(self.ctx, self.arrays[f].ptr, self.n * d, np.float32) for f, d in FIELDS}


def make_camera(cam: dict) -> _abi.Camera:
    c = _abi.Camera()
    c.width, c.height = int(cam["width"]), int(cam["height"])
    c.fx, c.fy, c.cx, c.cy = (float(np.float32(cam[k])) for k in ("fx", "fy", "cx", "cy"))
    c.R[:] = [float(v) for v in np.asarray(cam["R"], np.float32).reshape(9)]
    c.t[:] = [float(v) for v in np.asarray(cam["t"], np.float32).reshape(3)]
    return c


def make_settings(tile_size=16, near_plane=0.01, eps2d=0.3, background=(0.0, 0.0, 0.0),
                  min_alpha=1.0 / 255.0, min_transmittance=1e-4) -> _abi.RasterSettings:
    """RasterSettings defaults (Rasterizer.h:23-32)."""
    s = _abi.RasterSettings()
    s.tile_size = int(tile_size)
    s.near_plane = float(np.float32(near_plane))
    s.eps2d = float(np.float32(eps2d))
    s.background[:] = [float(np.float32(b)) for b in background]
    s.min_alpha = float(np.float32(min_alpha))
    s.min_transmittance = float(np.float32(min_transmittance))
    return s


class Rasterizer:
    """Rasterizer (Rasterizer.h:38-90) on the device."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.lib.u3d_raster_create(ctx.h, C.byref(h)))
        self.h = h
        self.n = 0
        ctx._adopt(self)

    def _release(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.u3d_raster_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def forward(self, model: DeviceGaussians, cam: dict, settings=None, out: DeviceArray | None = None,
                sh_rest: DeviceArray | None = None, sh_degree: int = 0):
        settings = settings or make_settings()
        c = make_camera(cam)
        if out is None:
            out = DeviceArray(self.ctx, (c.height * c.width * 3,))
        m = model.struct()
        if sh_rest is not None:
            check(self.ctx.lib.u3d_raster_forward_sh(self.h, C.byref(m), sh_rest.ptr, sh_degree,
                                                     C.byref(c), C.byref(settings), out.ptr))
        else:
            check(self.ctx.lib.u3d_raster_forward(self.h, C.byref(m), C.byref(c), C.byref(settings), out.ptr))
        self.n, self.cam = model.n, c
        return out

    def backward(self, model: DeviceGaussians, dl: DeviceArray, grads: DeviceGaussians,
                 mean2d_grad_ndc: DeviceArray | None = None, sh_rest: DeviceArray | None = None,
                 sh_degree: int = 0, grad_sh_rest: DeviceArray | None = None):
        m, g = model.struct(), grads.struct()
        ndc = mean2d_grad_ndc.ptr if mean2d_grad_ndc else None
        if sh_rest is not None:
            check(self.ctx.lib.u3d_raster_backward_sh(self.h, C.byref(m), sh_rest.ptr, sh_degree, dl.ptr,
                                                      C.byref(g), grad_sh_rest.ptr if grad_sh_rest else None,
                                                      ndc))
        else:
            check(self.ctx.lib.u3d_raster_backward(self.h, C.byref(m), dl.ptr, C.byref(g), ndc))
        return grads

    def radii(self) -> np.ndarray:
        p, n = C.c_void_p(), C.c_int64()
        check(self.ctx.lib.u3d_raster_radii(self.h, C.byref(p), C.byref(n)))
        return _view_numpy(self.ctx, p.value, n.value, np.int32)

    def state(self) -> dict:
        s = _abi.RasterState()
        check(self.ctx.lib.u3d_raster_get_state(self.h, C.byref(s)))
        n, K, tiles = s.n, s.num_pairs, s.tiles_x * s.tiles_y
        P = self.cam.width * self.cam.height
        v = lambda p, cnt, dt: _view_numpy(self.ctx, p, cnt, dt)
        return dict(n=n, num_pairs=K, tiles_x=s.tiles_x, tiles_y=s.tiles_y,
                    radius=v(s.radius, n, np.int32), rect=v(s.rect, 4 * n, np.int32),
                    depth=v(s.depth, n, np.float32), mean2d=v(s.mean2d, 2 * n, np.float32),
                    conic=v(s.conic, 3 * n, np.float32), rgb=v(s.color, 3 * n, np.float32),
                    alpha=v(s.opacity, n, np.float32), flags=v(s.flags, n, np.uint8),
                    tile_offsets=v(s.tile_offsets, tiles + 1, np.int32),
                    tile_list=v(s.tile_list, K, np.int32), final_t=v(s.final_t, P, np.float32),
                    n_contrib=v(s.n_contrib, P, np.int32))

    def grad2d(self) -> np.ndarray:
        p = C.c_void_p()
        check(self.ctx.lib.u3d_raster_get_grad2d(self.h, C.byref(p)))
        return _view_numpy(self.ctx, p.value, self.n * 9, np.float32)


def photometric_loss(ctx: Context, render: DeviceArray, target: DeviceArray, width: int,
                     height: int, ssim_lambda: float, grad: DeviceArray):
    """photometricLoss (Loss.h:18-19): returns (total, l1, ssim); overwrites grad."""
    v = _abi.LossValue()
    check(ctx.lib.u3d_photometric_loss(ctx.h, render.ptr, target.ptr, width, height,
                                       float(np.float32(ssim_lambda)), grad.ptr, C.byref(v)))
    return v.total, v.l1, v.ssim


def psnr(ctx: Context, a: DeviceArray, b: DeviceArray, count: int) -> float:
    out = C.c_float()
    check(ctx.lib.u3d_psnr(ctx.h, a.ptr, b.ptr, count, C.byref(out)))
    return out.value


def learning_rates(means=1.6e-4, scales=5e-3, quats=1e-3, opacities=5e-2, colors=2.5e-3):
    """LearningRates defaults (Adam.h:8-14)."""
    return _abi.LearningRates(means, scales, quats, opacities, colors)


class Adam:
    """Adam (Adam.h:18-33) with device moments."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.lib.u3d_adam_create(ctx.h, C.byref(h)))
        self.h = h
        ctx._adopt(self)

    def _release(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.u3d_adam_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def reset(self, n: int):
        check(self.ctx.lib.u3d_adam_reset(self.h, n))

    def compact(self, keep: DeviceArray):
        check(self.ctx.lib.u3d_adam_compact(self.h, keep.ptr))

    def grow(self, count: int):
        check(self.ctx.lib.u3d_adam_grow(self.h, count))

    def reset_opacity_moments(self):
        check(self.ctx.lib.u3d_adam_reset_opacity_moments(self.h))

    def step(self, params: DeviceGaussians, grads: DeviceGaussians, lr):
        p, g = params.struct(), grads.struct()
        check(self.ctx.lib.u3d_adam_step(self.h, C.byref(p), C.byref(g), C.byref(lr)))

    def moments(self):
        m, v, st = _abi.Gaussians(), _abi.Gaussians(), C.c_int()
        check(self.ctx.lib.u3d_adam_moments(self.h, C.byref(m), C.byref(v), C.byref(st)))
        out = []
        for g in (m, v):
            out.append({f: _view_numpy(self.ctx, getattr(g, f), g.n * d, np.float32) for f, d in FIELDS})
        return out[0], out[1], st.value


def densify_decide(ctx: Context, model: DeviceGaussians, accum: DeviceArray, count: DeviceArray,
                   extent: float, max_gaussians: int, grad_threshold=2e-4, percent_dense=0.01,
                   prune_opacity=0.005, opacity_was_reset=False):
    """Decisions of Trainer::densifyAndPrune: (keep, kind) device arrays and host counts."""
    n = model.n
    keep = DeviceArray(ctx, (max(n, 1),), np.uint8)
    kind = DeviceArray(ctx, (max(n, 1),), np.uint8)
    nc, ns, nk = C.c_int64(), C.c_int64(), C.c_int64()
    m = model.struct()
    check(ctx.lib.u3d_densify_decide(ctx.h, C.byref(m), accum.ptr, count.ptr, float(extent),
                                     int(max_gaussians), float(grad_threshold), float(percent_dense),
                                     float(prune_opacity), int(bool(opacity_was_reset)), keep.ptr,
                                     kind.ptr, C.byref(nc), C.byref(ns), C.byref(nk)))
    return keep, kind, dict(clones=nc.value, splits=ns.value, kept=nk.value)


def densify_apply(ctx: Context, model: DeviceGaussians, keep: DeviceArray, kind: DeviceArray,
                  normals: DeviceArray | None, log_shrink: float) -> int:
    m = model.struct()
    nn = C.c_int64()
    check(ctx.lib.u3d_densify_apply(ctx.h, C.byref(m), keep.ptr, kind.ptr,
                                    normals.ptr if normals is not None else None,
                                    float(log_shrink), C.byref(nn)))
    model.n = nn.value
    return nn.value


def options_for_iterations(iterations: int) -> _abi.TrainOptions:
    """TrainOptions defaults (Trainer.h:19-37) with TrainOptions::forIterations (Trainer.cpp:26-35)."""
    o = _abi.TrainOptions()
    o.iterations = max(1, int(iterations))
    o.max_gaussians = 200000
    o.ssim_lambda = 0.2
    o.lr = learning_rates()
    o.means_lr_final = 1.6e-6
    o.densify_grad_threshold = 2e-4
    o.percent_dense = 0.01
    o.refine_start = min(500, o.iterations // 10)
    o.refine_stop = o.iterations * 6 // 10
    o.refine_interval = 100
    o.opacity_reset_interval = 3000 if o.iterations >= 6000 else 0
    o.prune_opacity = 0.005
    o.seed = 42
    o.raster = make_settings()
    return o


class Trainer:
    """Trainer (Trainer.h:249-292) with all state on the device."""

    def __init__(self, ctx: Context, initial, cams: list, images_u8: list, extent: float,
                 options: _abi.TrainOptions):
        self.ctx = ctx
        get = (lambda f: initial[f]) if isinstance(initial, dict) else (lambda f: getattr(initial, f))
        arrs = [np.ascontiguousarray(get(f), np.float32).reshape(-1) for f, _ in FIELDS]
        n = len(arrs[3])
        cam_arr = (_abi.Camera * max(len(cams), 1))(*[make_camera(c) for c in cams])
        rgb = np.ascontiguousarray(np.concatenate([np.asarray(i, np.uint8).reshape(-1) for i in images_u8])
                                   if images_u8 else np.zeros(1, np.uint8))
        h = C.c_void_p()
        check(ctx.lib.u3d_trainer_create(ctx.h, *(a.ctypes.data for a in arrs), n, cam_arr,
                                         rgb.ctypes.data, len(cams), float(extent),
                                         C.byref(options), C.byref(h)))
        self.h = h
        ctx._adopt(self)

    def _release(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.u3d_trainer_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def step(self):
        check(self.ctx.lib.u3d_trainer_step(self.h))

    def done(self) -> bool:
        return bool(self.ctx.lib.u3d_trainer_done(self.h))

    def iteration(self) -> int:
        return int(self.ctx.lib.u3d_trainer_iteration(self.h))

    def size(self) -> int:
        return int(self.ctx.lib.u3d_trainer_size(self.h))

    def model(self) -> dict:
        n = self.size()
        out = {f: np.zeros(n * d, np.float32) for f, d in FIELDS}
        check(self.ctx.lib.u3d_trainer_get_model(self.h, *(out[f].ctypes.data for f, _ in FIELDS)))
        return out

    def moments(self) -> tuple:
        """Adam moments (m, v), each a dict of the model fields."""
        n = self.size()
        m = np.zeros(14 * n, np.float32)
        v = np.zeros(14 * n, np.float32)
        check(self.ctx.lib.u3d_trainer_get_moments(self.h, m.ctypes.data, v.ctypes.data))
        return _split_fields(m, n), _split_fields(v, n)

    def counters(self) -> dict:
        c = (C.c_int64 * _abi.U3D_NUM_COUNTERS)()
        check(self.ctx.lib.u3d_trainer_counters(self.h, c))
        return dict(zip(_abi.COUNTERS, list(c)))

    def tuning(self) -> dict:
        t = _abi.BatchTuning()
        check(self.ctx.lib.u3d_trainer_get_tuning(self.h, C.byref(t)))
        return _tuning_dict(t)

    def set_tuning(self, heavy_first=1, workgroups_per_cu=0, lds_batch=256, sort_items=0, tile_size=0,
                   atomic_lanes=0, sort_radix=0, pair_sort=0):
        t = _abi.BatchTuning(int(heavy_first), int(workgroups_per_cu), int(lds_batch), int(sort_items),
                             int(tile_size), int(atomic_lanes), int(sort_radix), int(pair_sort))
        check(self.ctx.lib.u3d_trainer_set_tuning(self.h, C.byref(t)))

    def autotune(self, steps_per_candidate: int = 3, tune_tile_size: bool = False) -> dict:
        t = _abi.BatchTuning()
        check(self.ctx.lib.u3d_trainer_autotune(self.h, int(steps_per_candidate), int(tune_tile_size),
                                                C.byref(t)))
        return _tuning_dict(t)

    def export_ply(self, path: str):
        check(self.ctx.lib.u3d_trainer_export_ply(self.h, path.encode()))

    def sh_rest(self) -> np.ndarray:
        """SH coefficients [size x 15 x 3] (trainers created with sh_degree > 0)."""
        out = np.zeros(self.size() * 45, np.float32)
        check(self.ctx.lib.u3d_trainer_get_sh_rest(self.h, out.ctypes.data))
        return out

    def set_sh_rest(self, rest: np.ndarray):
        a = np.ascontiguousarray(rest, np.float32).reshape(-1)
        assert a.size == self.size() * 45
        check(self.ctx.lib.u3d_trainer_set_sh_rest(self.h, a.ctypes.data))

    def sh_moments(self) -> tuple:
        """Adam moments (m, v) of the SH coefficients, [size x 15 x 3] each."""
        m = np.zeros(self.size() * 45, np.float32)
        v = np.zeros(self.size() * 45, np.float32)
        check(self.ctx.lib.u3d_trainer_get_sh_moments(self.h, m.ctypes.data, v.ctypes.data))
        return m, v

    def densify_stats(self) -> tuple:
        """(accumulated NDC mean-gradient norms, visible-step counts) since the last densify."""
        n = self.size()
        accum = np.zeros(n, np.float32)
        count = np.zeros(n, np.int32)
        check(self.ctx.lib.u3d_trainer_get_densify_stats(self.h, accum.ctypes.data, count.ctypes.data))
        return accum, count

    def stats(self) -> dict:
        s = _abi.TrainStats()
        check(self.ctx.lib.u3d_trainer_stats(self.h, C.byref(s)))
        return dict(iteration=s.iteration, loss=s.loss, psnr=s.psnr, gaussians=s.gaussians)

    def evaluate_psnr(self) -> float:
        v = C.c_float()
        check(self.ctx.lib.u3d_trainer_evaluate_psnr(self.h, C.byref(v)))
        return v.value

    def evaluate_l1(self) -> float:
        v = C.c_float()
        check(self.ctx.lib.u3d_trainer_evaluate_l1(self.h, C.byref(v)))
        return v.value

    def render(self, cam: dict) -> np.ndarray:
        c = make_camera(cam)
        out = np.zeros(c.width * c.height * 3, np.float32)
        check(self.ctx.lib.u3d_trainer_render(self.h, C.byref(c), out.ctypes.data))
        return out

    def synchronize(self):
        self.ctx.synchronize()


def _tuning_dict(t) -> dict:
    return dict(heavy_first=t.heavy_first, workgroups_per_cu=t.workgroups_per_cu,
                lds_batch=t.lds_batch, sort_items=t.sort_items, tile_size=t.tile_size,
                atomic_lanes=t.atomic_lanes, sort_radix=t.sort_radix, pair_sort=t.pair_sort)


def _split_fields(flat: np.ndarray, n: int) -> dict:
    out, o = {}, 0
    for f, d in FIELDS:
        out[f] = flat[o:o + d * n]
        o += d * n
    return out


def _cams_array(cams):
    if not cams:
        return None, 0
    arr = (_abi.Camera * len(cams))(*[make_camera(c) for c in cams])
    return arr, len(cams)


def points_knn_init(ctx: Context, xyz: np.ndarray, rgb: np.ndarray | None = None) -> dict:
    """initGaussiansFromPoints (GaussianModel.cpp:68-144) on the device; returns host arrays."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    n = len(xyz) // 3
    rgb = np.zeros(0, np.float32) if rgb is None else np.ascontiguousarray(rgb, np.float32).reshape(-1)
    dx = DeviceArray.from_numpy(ctx, xyz)
    dr = DeviceArray.from_numpy(ctx, rgb if len(rgb) else np.zeros(1, np.float32))
    out = DeviceGaussians(ctx, n)
    m = out.struct()
    check(ctx.lib.u3d_points_knn_init(ctx.h, dx.ptr, dr.ptr, len(rgb), n, C.byref(m)))
    return {k: v.copy() for k, v in out.to_host().items()}


def points_features(ctx: Context, xyz, rgb, cams=None, near_plane: float = 0.01) -> np.ndarray:
    """Learned-initialiser inputs [n x 80] (host copy)."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1)
    n = len(xyz) // 3
    dx, dr = DeviceArray.from_numpy(ctx, xyz), DeviceArray.from_numpy(ctx, rgb)
    f = DeviceArray(ctx, (max(n, 1) * _abi.FEATURE_DIM,))
    arr, nc = _cams_array(cams)
    check(ctx.lib.u3d_points_features(ctx.h, dx.ptr, dr.ptr, len(rgb), n, arr, nc,
                                      float(near_plane), f.ptr))
    return f.numpy()[:n * _abi.FEATURE_DIM].reshape(n, _abi.FEATURE_DIM)


def write_splat_ply(path: str, model: dict, sh_rest=None, sh_degree: int = 0):
    """writeSplatPly (SplatPly.cpp:261-281) with optional f_rest_* SH coefficients."""
    lib = _abi.load()
    arrs = [np.ascontiguousarray(model[f], np.float32).reshape(-1) for f, _ in FIELDS]
    n = len(arrs[3])
    sh = None if sh_rest is None else np.ascontiguousarray(sh_rest, np.float32).reshape(-1)
    check(lib.u3d_write_splat_ply(path.encode(), *(a.ctypes.data for a in arrs), n,
                                  sh.ctypes.data if sh is not None else None, int(sh_degree)))


def read_splat_ply(path: str):
    """readSplatPly for 3DGS files: (model dict, sh_rest [n*45] or None, sh_degree)."""
    lib = _abi.load()
    n, deg = C.c_int64(), C.c_int()
    check(lib.u3d_read_splat_ply(path.encode(), 0, None, None, None, None, None, None,
                                 C.byref(n), C.byref(deg)))
    out = {f: np.zeros(max(n.value * d, 1), np.float32) for f, d in FIELDS}
    sh = np.zeros(max(n.value * 45, 1), np.float32)
    check(lib.u3d_read_splat_ply(path.encode(), max(n.value, 1), *(out[f].ctypes.data for f, _ in FIELDS),
                                 sh.ctypes.data, C.byref(n), C.byref(deg)))
    model = {f: out[f][:n.value * d] for f, d in FIELDS}
    return model, (sh[:n.value * 45] if deg.value > 0 else None), deg.value


def ply_info(path: str) -> dict:
    """Vertex count, whether the file is a Gaussian splat file, and its SH degree."""
    lib = _abi.load()
    n, splat, deg = C.c_int64(), C.c_int(), C.c_int()
    check(lib.u3d_ply_info(path.encode(), C.byref(n), C.byref(splat), C.byref(deg)))
    return dict(vertices=n.value, is_splat=bool(splat.value), sh_degree=deg.value)


def write_points_ply(path: str, xyz, rgb=None):
    """writePointsPly (SplatPly.cpp:344-362): rgb in 0..1 stored as uchar."""
    lib = _abi.load()
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    rgb = np.zeros(0, np.float32) if rgb is None else np.ascontiguousarray(rgb, np.float32).reshape(-1)
    check(lib.u3d_write_points_ply(path.encode(), xyz.ctypes.data, rgb.ctypes.data if len(rgb) else None,
                                   len(rgb), len(xyz) // 3))


def read_points_ply(path: str):
    """readPointsPly (SplatPly.cpp:364-386): (xyz [3n], rgb [3n] in 0..1)."""
    lib = _abi.load()
    n = C.c_int64()
    check(lib.u3d_read_points_ply(path.encode(), 0, None, None, C.byref(n)))
    xyz = np.zeros(max(3 * n.value, 1), np.float32)
    rgb = np.zeros(max(3 * n.value, 1), np.float32)
    check(lib.u3d_read_points_ply(path.encode(), max(n.value, 1), xyz.ctypes.data, rgb.ctypes.data,
                                  C.byref(n)))
    return xyz[:3 * n.value], rgb[:3 * n.value]


def view_attributes(ctx: Context, model) -> tuple:
    """loadSplats' display attributes (SplatView.cpp:117-180) on the device: (rgba [4n] uint8,
    cov6 [6n] float32) as host copies."""
    d = model if isinstance(model, DeviceGaussians) else DeviceGaussians.from_host(ctx, model)
    rgba = DeviceArray(ctx, (max(d.n, 1),), np.uint32)
    cov = DeviceArray(ctx, (max(6 * d.n, 1),))
    m = d.struct()
    check(ctx.lib.u3d_view_splat_attributes(ctx.h, C.byref(m), rgba.ptr, cov.ptr))
    return (_view_numpy(ctx, rgba.ptr, 4 * d.n, np.uint8),
            _view_numpy(ctx, cov.ptr, 6 * d.n, np.float32))


def view_depth_order(ctx: Context, centers, view_row2, rgba=None, cov6=None):
    """sortAndUpload (SplatView.cpp:349-397) on the device: the front-to-back order [n] and,
    given rgba and cov6, the packed 40-byte records as a structured array."""
    c = np.ascontiguousarray(centers, np.float32).reshape(-1)
    n = len(c) // 3
    row = np.ascontiguousarray(view_row2, np.float32).reshape(4)
    dc = DeviceArray.from_numpy(ctx, c if n else np.zeros(3, np.float32))
    order = DeviceArray(ctx, (max(n, 1),), np.uint32)
    pack = rgba is not None
    if pack:
        drgba = DeviceArray.from_numpy(ctx, np.ascontiguousarray(rgba, np.uint8).reshape(-1).view(np.uint32)
                                       if n else np.zeros(1, np.uint32))
        dcov = DeviceArray.from_numpy(ctx, np.ascontiguousarray(cov6, np.float32).reshape(-1)
                                      if n else np.zeros(1, np.float32))
        packed = DeviceArray(ctx, (max(10 * n, 1),), np.float32)
    check(ctx.lib.u3d_view_depth_order(ctx.h, dc.ptr, n, row.ctypes.data, order.ptr,
                                       drgba.ptr if pack else None, dcov.ptr if pack else None,
                                       packed.ptr if pack else None))
    o = _view_numpy(ctx, order.ptr, n, np.uint32)
    if not pack:
        return o
    rec = np.dtype([("center", np.float32, 3), ("color", np.uint8, 4), ("cov", np.float32, 6)])
    return o, _view_numpy(ctx, packed.ptr, 40 * n, np.uint8).view(rec)


class InitNet:
    """The learned initialiser's network on the device (u3d_init_net)."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.lib.u3d_init_net_create(ctx.h, int(seed), C.byref(h)))
        self.h = h
        self.count = int(ctx.lib.u3d_init_net_param_count())
        ctx._adopt(self)

    def _release(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.u3d_init_net_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.count, np.float32)
        check(self.ctx.lib.u3d_init_net_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.count
        check(self.ctx.lib.u3d_init_net_set_params(self.h, p.ctypes.data))

    def grads(self) -> np.ndarray:
        out = np.zeros(self.count, np.float32)
        check(self.ctx.lib.u3d_init_net_get_grads(self.h, out.ctypes.data))
        return out

    def forward(self, features: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, _abi.FEATURE_DIM)
        n = len(f)
        df = DeviceArray.from_numpy(self.ctx, f)
        out = DeviceArray(self.ctx, (max(n, 1) * _abi.NET_OUT,))
        check(self.ctx.lib.u3d_init_net_forward(self.h, df.ptr, n, out.ptr))
        return out.numpy()[:n * _abi.NET_OUT].reshape(n, _abi.NET_OUT)

    def train_step(self, features, targets, weights, lr: float) -> float:
        """features / targets: DeviceArray (n x 80, n x 16) or numpy arrays."""
        if isinstance(features, np.ndarray):
            features = DeviceArray.from_numpy(self.ctx, np.ascontiguousarray(features, np.float32))
        if isinstance(targets, np.ndarray):
            targets = DeviceArray.from_numpy(self.ctx, np.ascontiguousarray(targets, np.float32))
        n = features.nbytes // (4 * _abi.FEATURE_DIM)
        w = np.ascontiguousarray(weights, np.float32).reshape(16)
        loss = C.c_float()
        check(self.ctx.lib.u3d_init_net_train_step(self.h, features.ptr, targets.ptr, w.ctypes.data,
                                                   n, float(lr), C.byref(loss)))
        return loss.value


def learned_init(ctx: Context, net: "InitNet | None", xyz, rgb, cams=None,
                 near_plane: float = 0.01) -> dict:
    """kNN initialisation corrected by the network (net None: kNN only); host arrays."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1)
    n = len(xyz) // 3
    dx, dr = DeviceArray.from_numpy(ctx, xyz), DeviceArray.from_numpy(ctx, rgb)
    out = DeviceGaussians(ctx, n)
    m = out.struct()
    arr, nc = _cams_array(cams)
    check(ctx.lib.u3d_points_learned_init(ctx.h, net.h if net else None, dx.ptr, dr.ptr, len(rgb), n,
                                          arr, nc, float(near_plane), C.byref(m)))
    return {k: v.copy() for k, v in out.to_host().items()}


class _Member(Trainer):
    """Borrowed handle of a scene-set member: the set owns (and destroys) the C object, and the
    handle is valid while the set lives."""

    def _release(self):
        self.h = None

    def __del__(self):
        pass

    def step(self):
        raise RuntimeError("members of a scene set are stepped through the set")


class SceneSet:
    """Independent scenes trained together by the batched engine (u3d_scene_set)."""

    def __init__(self, ctx: Context, use_graph: bool = True, hbm_budget_bytes: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.lib.u3d_scene_set_create(ctx.h, int(bool(use_graph)), int(hbm_budget_bytes),
                                           C.byref(h)))
        self.h = h
        self._members = []
        ctx._adopt(self)

    def _release(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.lib.u3d_scene_set_destroy(self.h)
        self.h = None
        for m in self._members:
            m.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def add(self, trainer: Trainer) -> bool:
        """Admits the trainer when it fits the HBM budget (the set then owns it)."""
        rc = self.ctx.lib.u3d_scene_set_add(self.h, trainer.h)
        if rc == 3:  # U3D_ERR_CAPACITY
            return False
        check(rc)
        trainer.__class__ = _Member  # the set owns the C object now (no back reference: no cycle)
        self._members.append(trainer)
        return True

    def __len__(self):
        return int(self.ctx.lib.u3d_scene_set_size(self.h))

    def member(self, k: int) -> Trainer:
        return self._members[k]

    def step(self):
        check(self.ctx.lib.u3d_scene_set_step(self.h))

    def done(self) -> bool:
        return bool(self.ctx.lib.u3d_scene_set_done(self.h))

    def bytes(self) -> int:
        return int(self.ctx.lib.u3d_scene_set_bytes(self.h))

    def set_tuning(self, heavy_first=1, workgroups_per_cu=0, lds_batch=256, sort_items=0, tile_size=0,
                   atomic_lanes=0, sort_radix=0, pair_sort=0):
        t = _abi.BatchTuning(int(heavy_first), int(workgroups_per_cu), int(lds_batch), int(sort_items),
                             int(tile_size), int(atomic_lanes), int(sort_radix), int(pair_sort))
        check(self.ctx.lib.u3d_scene_set_set_tuning(self.h, C.byref(t)))

    def autotune(self, steps_per_candidate: int = 3, tune_tile_size: bool = False) -> dict:
        t = _abi.BatchTuning()
        check(self.ctx.lib.u3d_scene_set_autotune(self.h, int(steps_per_candidate),
                                                  int(tune_tile_size), C.byref(t)))
        return _tuning_dict(t)
