"""Seeded synthetic scenes for the Gaussian Splatting trainer (benchmark and tests).

All randomness comes from numpy's counter-based Philox generator, so scenes are identical on
every host. Distributions follow the workloads the trainer is measured on:

* ``random_gaussians``: means uniform in a ball, log-scales N(mu, sigma), uniformly random unit
  quaternions, opacity logits N(0, 1), SH-DC colours from an RGB uniform in [0, 1].
* ``ring_cameras``: pinhole cameras on a ring around the origin at alternating +-elevation,
  looking at the origin (OpenCV convention: x right, y down, z forward), fx = fy = 0.8 W.
* ``hemisphere_cameras``: views spread over the upper hemisphere (Fibonacci spiral).
* ``surface_points``: noisy points on a sphere-and-torus union with random RGB.

Camera dicts carry width, height, fx, fy, cx, cy, R (row-major world->camera) and t, the
fields of the reference's RenderCamera (engine/include/u3d/splat/Rasterizer.h:13-21).
"""
from __future__ import annotations

import math

import numpy as np

SH_C0 = np.float32(0.28209479177387814)  # engine/include/u3d/splat/GaussianModel.h:10


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def random_gaussians(n: int, seed: int, radius: float = 1.0, log_scale_mean: float = -4.0,
                     log_scale_std: float = 0.5, opacity_std: float = 1.0):
    """SoA dict (means, scales, quats, opacities, colors) as float32 flat arrays."""
    g = rng(seed)
    d = g.normal(size=(n, 3))
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-12)
    r = radius * np.cbrt(g.random(n))
    means = d * r[:, None]
    scales = g.normal(log_scale_mean, log_scale_std, size=(n, 3))
    q = g.normal(size=(n, 4))
    q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    opac = g.normal(0.0, opacity_std, size=n)
    rgb = g.random((n, 3))
    colors = (rgb - 0.5) / float(SH_C0)
    f = lambda a: np.ascontiguousarray(a, np.float32).reshape(-1)
    return dict(means=f(means), scales=f(scales), quats=f(q), opacities=f(opac), colors=f(colors))


def look_at(center, target, width: int, height: int, focal: float):
    """Camera at `center` looking at `target`, image y along world +y (down), OpenCV axes."""
    C = np.asarray(center, np.float64)
    z = np.asarray(target, np.float64) - C
    z /= np.linalg.norm(z)
    down = np.array([0.0, 1.0, 0.0])
    x = np.cross(down, z)
    if np.linalg.norm(x) < 1e-9:
        x = np.cross(np.array([0.0, 0.0, 1.0]), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    t = -R @ C
    return dict(width=int(width), height=int(height), fx=float(focal), fy=float(focal),
                cx=0.5 * (width - 1), cy=0.5 * (height - 1),
                R=R.astype(np.float32).reshape(9), t=t.astype(np.float32))


def ring_cameras(count: int, width: int, height: int, radius: float = 3.0,
                 elevation_deg: float = 15.0, focal_frac: float = 0.8):
    cams = []
    for k in range(count):
        a = 2.0 * math.pi * k / count
        e = math.radians(elevation_deg) * (1.0 if k % 2 == 0 else -1.0)
        C = (radius * math.cos(e) * math.sin(a), -radius * math.sin(e), -radius * math.cos(e) * math.cos(a))
        cams.append(look_at(C, (0.0, 0.0, 0.0), width, height, focal_frac * width))
    return cams


def hemisphere_cameras(count: int, width: int, height: int, radius: float = 3.0,
                       focal_frac: float = 0.8):
    """Fibonacci spiral over the upper hemisphere (world up is -y)."""
    cams = []
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for k in range(count):
        h = 0.1 + 0.85 * (k + 0.5) / count  # height fraction above the equator
        r = math.sqrt(max(0.0, 1.0 - h * h))
        a = golden * k
        C = (radius * r * math.cos(a), -radius * h, radius * r * math.sin(a))
        cams.append(look_at(C, (0.0, 0.0, 0.0), width, height, focal_frac * width))
    return cams


def orbit_camera(w: int, h: int, f: float, angle: float, radius: float, height: float):
    """Camera on a circle around the origin looking at it (world y down)."""
    C = (radius * math.sin(angle), height, -radius * math.cos(angle))
    return look_at(C, (0.0, 0.0, 0.0), w, h, f)


def scene_extent(cams) -> float:
    """1.1 x the largest camera distance from the mean camera centre
    (engine/src/pipeline/Reconstructor.cpp:294-304)."""
    centres = []
    for c in cams:
        R = np.asarray(c["R"], np.float64).reshape(3, 3)
        t = np.asarray(c["t"], np.float64)
        centres.append(-R.T @ t)
    centres = np.array(centres)
    m = centres.mean(axis=0)
    return max(float(np.max(np.linalg.norm(centres - m, axis=1))) * 1.1, 1e-3)


def surface_points(n: int, seed: int, noise: float = 0.01):
    """Points on the union of a sphere (r = 0.6) and a torus (R = 0.8, r = 0.2), plus noise."""
    g = rng(seed)
    n_s = n // 2
    d = g.normal(size=(n_s, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    sphere = 0.6 * d
    u = g.random(n - n_s) * 2 * math.pi
    v = g.random(n - n_s) * 2 * math.pi
    torus = np.stack([(0.8 + 0.2 * np.cos(v)) * np.cos(u), 0.2 * np.sin(v),
                      (0.8 + 0.2 * np.cos(v)) * np.sin(u)], axis=1)
    pts = np.concatenate([sphere, torus]) + g.normal(0.0, noise, size=(n, 3))
    rgb = g.random((n, 3))
    return (np.ascontiguousarray(pts, np.float32).reshape(-1),
            np.ascontiguousarray(rgb, np.float32).reshape(-1))


def quantize(rgb01) -> np.ndarray:
    """TrainImage::quantize (engine/src/splat/Trainer.cpp:10-16): round(clamp(x) * 255)."""
    x = np.clip(np.asarray(rgb01, np.float32), np.float32(0), np.float32(1)) * np.float32(255)
    # std::lround rounds halves away from zero; x >= 0 here.
    return np.floor(x.astype(np.float64) + 0.5).astype(np.uint8)
