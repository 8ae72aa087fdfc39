"""Learned initialiser: training data, network training and the time-to-target-L1 comparison.

No reference counterpart (the reference initialises every seed point with the kNN-distance
heuristic, GaussianModel.cpp:68-144, which stays the fallback here).

Scenes are unions of spheres, tori and boxes. Their ground truth is a set of flat Gaussians
lying on the surface (oriented with it, coloured by a smooth procedural field); seed points are
other noisy samples of the same surface carrying the field's colour, like SfM / AR points.
The plain trainer converges each training scene from the kNN initialisation of its seeds
(no densification, so Gaussian i stays seed i); the network then learns, per seed point, the
difference between the converged parameters and the kNN initialisation, from the point's
neighbourhood features. At inference the correction is added to the kNN initialisation.

Targets put each converged Gaussian in a canonical frame (axes by decreasing scale, right
handed, the largest component of the first two axes positive, quaternion w >= 0), so that
equal covariances give equal targets.

Run as a module to build the dataset, train the network on the device and evaluate:
    python -m univeral_3d_reconstruction_system_amd.learned_init --help
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np

from . import _abi
from . import splat as U
from . import synthetic as S

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "init_net.npz")
OUTPUT_WEIGHTS = np.array([1, 1, 1, 1, 1, 1, 1, 0.25, 1, 1, 1, 0, 0, 0, 0, 0], np.float32)
LOGIT_01 = float(np.log(np.float32(0.1) / (np.float32(1) - np.float32(0.1))))


# ------------------------------------------------------------------------------ scenes
def _field(p, phase):
    """Smooth procedural colour field in [0.05, 0.95]."""
    c = np.stack([np.sin(3.1 * p[:, 0] + phase[0]) * np.cos(2.3 * p[:, 1] + phase[1]),
                  np.sin(2.7 * p[:, 1] + phase[2]) * np.cos(3.3 * p[:, 2] + phase[3]),
                  np.sin(2.9 * p[:, 2] + phase[4]) * np.cos(2.1 * p[:, 0] + phase[5])], axis=1)
    return 0.5 + 0.45 * c


def _sample_primitive(g, kind, n, centre, size, rot):
    """n surface points and unit normals of a primitive (world frame)."""
    if kind == 0:  # sphere
        d = g.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        p, nrm = size * d, d
    elif kind == 1:  # torus, tube radius 0.3 size
        R, r = size, 0.3 * size
        u, v = g.random(n) * 2 * math.pi, g.random(n) * 2 * math.pi
        ring = np.stack([np.cos(u), np.zeros(n), np.sin(u)], axis=1)
        nrm = np.cos(v)[:, None] * ring + np.sin(v)[:, None] * np.array([0.0, 1.0, 0.0])
        p = R * ring + r * nrm
    else:  # box with half-extent size
        face = g.integers(0, 6, n)
        uv = g.uniform(-1, 1, size=(n, 2)) * size
        ax = face // 2
        sgn = np.where(face % 2 == 0, 1.0, -1.0)
        p = np.zeros((n, 3))
        nrm = np.zeros((n, 3))
        for a in range(3):
            m = ax == a
            o = [b for b in range(3) if b != a]
            p[m, a] = sgn[m] * size
            p[m, o[0]] = uv[m, 0]
            p[m, o[1]] = uv[m, 1]
            nrm[m, a] = sgn[m]
    return p @ rot.T + centre, nrm @ rot.T


def _random_rotation(g):
    q = g.normal(size=4)
    q /= np.linalg.norm(q)
    return _quat_to_rot(q[None])[0]


def _quat_to_rot(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def _rot_to_quat(R):
    """Quaternions (w, x, y, z), w >= 0, of rotation matrices [n, 3, 3]."""
    m = R
    t = m[:, 0, 0] + m[:, 1, 1] + m[:, 2, 2]
    q = np.zeros((len(R), 4))
    for i in range(len(R)):
        a = m[i]
        if t[i] > 0:
            s = math.sqrt(t[i] + 1.0) * 2
            q[i] = [0.25 * s, (a[2, 1] - a[1, 2]) / s, (a[0, 2] - a[2, 0]) / s, (a[1, 0] - a[0, 1]) / s]
        elif a[0, 0] > a[1, 1] and a[0, 0] > a[2, 2]:
            s = math.sqrt(1.0 + a[0, 0] - a[1, 1] - a[2, 2]) * 2
            q[i] = [(a[2, 1] - a[1, 2]) / s, 0.25 * s, (a[0, 1] + a[1, 0]) / s, (a[0, 2] + a[2, 0]) / s]
        elif a[1, 1] > a[2, 2]:
            s = math.sqrt(1.0 + a[1, 1] - a[0, 0] - a[2, 2]) * 2
            q[i] = [(a[0, 2] - a[2, 0]) / s, (a[0, 1] + a[1, 0]) / s, 0.25 * s, (a[1, 2] + a[2, 1]) / s]
        else:
            s = math.sqrt(1.0 + a[2, 2] - a[0, 0] - a[1, 1]) * 2
            q[i] = [(a[1, 0] - a[0, 1]) / s, (a[0, 2] + a[2, 0]) / s, (a[1, 2] + a[2, 1]) / s, 0.25 * s]
    q *= np.where(q[:, :1] < 0, -1.0, 1.0)
    return q


def mesh_scene(seed: int, n_gt: int, n_seed: int, views: int, width: int, height: int,
               noise: float = 0.01) -> dict:
    """Ground-truth Gaussians, seed points and cameras of one synthetic scene."""
    g = S.rng(10_000 + seed)
    k = int(g.integers(1, 4))
    kinds = g.integers(0, 3, k)
    centres = g.uniform(-0.35, 0.35, size=(k, 3))
    sizes = g.uniform(0.25, 0.45, size=k)
    rots = [_random_rotation(g) for _ in range(k)]
    phase = g.uniform(0, 2 * math.pi, size=6)
    share = np.maximum(sizes ** 2, 1e-3)
    share /= share.sum()

    def sample(n):
        cnt = np.maximum(1, np.round(share * n).astype(int))
        cnt[-1] = max(1, n - cnt[:-1].sum())
        ps, ns = zip(*[_sample_primitive(g, kinds[i], cnt[i], centres[i], sizes[i], rots[i])
                       for i in range(k)])
        return np.concatenate(ps)[:n], np.concatenate(ns)[:n]

    p, nrm = sample(n_gt)
    # flat Gaussians in the tangent plane: tangent scale from the sampling density
    tan = math.sqrt(4.0 * float(np.sum(share * 4 * math.pi * sizes ** 2)) / n_gt)
    a = np.cross(nrm, np.array([0.3, 0.9, 0.2]))
    a /= np.maximum(np.linalg.norm(a, axis=1, keepdims=True), 1e-9)
    b = np.cross(nrm, a)
    R = np.stack([a, b, nrm], axis=2)  # columns: tangent, tangent, normal
    det = np.linalg.det(R)
    R[det < 0, :, 1] *= -1
    quats = _rot_to_quat(R)
    ls = np.log(tan) + g.normal(0, 0.15, size=(n_gt, 2))
    scales = np.concatenate([ls, np.log(tan / 8.0) * np.ones((n_gt, 1))], axis=1)
    rgb = _field(p, phase)
    f = lambda x: np.ascontiguousarray(x, np.float32).reshape(-1)
    gt = dict(means=f(p), scales=f(scales), quats=f(quats),
              opacities=f(np.full(n_gt, 2.2)), colors=f((rgb - 0.5) / S.SH_C0))
    sp, _ = sample(n_seed)
    sp = sp + g.normal(0, noise, size=sp.shape)
    srgb = np.clip(_field(sp, phase) + g.normal(0, 0.03, size=sp.shape), 0, 1)
    cams = S.ring_cameras(views, width, height)
    return dict(gt=gt, xyz=f(sp), rgb=f(srgb), cams=cams, extent=S.scene_extent(cams))


def render_targets(ctx: U.Context, gt: dict, cams) -> list:
    d = U.DeviceGaussians.from_host(ctx, gt)
    r = U.Rasterizer(ctx)
    out = U.DeviceArray(ctx, (cams[0]["width"] * cams[0]["height"] * 3,))
    return [S.quantize(r.forward(d, c, out=out).numpy()) for c in cams]


# ------------------------------------------------------------------- training data
def converge(ctx, init: dict, cams, images, extent: float, iterations: int) -> dict:
    """The plain trainer from `init`, no densification or opacity reset (Gaussian i = seed i)."""
    opt = U.options_for_iterations(iterations)
    opt.refine_start = 10 ** 9
    opt.opacity_reset_interval = 0
    opt.max_gaussians = len(init["opacities"])
    t = U.Trainer(ctx, init, cams, images, extent, opt)
    while not t.done():
        t.step()
    return t.model()


def canonical(scales, quats):
    """Canonical (log-scales, quaternion) of Gaussians: axes by decreasing scale, right-handed,
    largest component of the first two axes positive, w >= 0."""
    s = np.asarray(scales, np.float64).reshape(-1, 3)
    q = np.asarray(quats, np.float64).reshape(-1, 4)
    q = q / np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    R = _quat_to_rot(q)
    order = np.argsort(-s, axis=1)
    s2 = np.take_along_axis(s, order, axis=1)
    R2 = np.take_along_axis(R, order[:, None, :].repeat(3, 1), axis=2)
    for c in (0, 1):
        col = R2[:, :, c]
        big = col[np.arange(len(col)), np.argmax(np.abs(col), axis=1)]
        flip = big < 0
        R2[flip, :, c] *= -1
    det = np.linalg.det(R2)
    R2[det < 0, :, 2] *= -1
    return s2, _rot_to_quat(R2)


def targets_from(init: dict, conv: dict) -> np.ndarray:
    """Per-seed network targets [n x 16]: converged minus kNN-initialised parameters."""
    n = len(init["opacities"])
    s, q = canonical(conv["scales"], conv["quats"])
    t = np.zeros((n, 16), np.float32)
    t[:, 0:3] = np.clip(s - np.asarray(init["scales"]).reshape(-1, 3), -3, 3)
    t[:, 3:7] = q - np.array([1.0, 0.0, 0.0, 0.0])
    t[:, 7] = np.clip(np.asarray(conv["opacities"]) - np.asarray(init["opacities"]), -8, 8)
    t[:, 8:11] = np.clip(np.asarray(conv["colors"]).reshape(-1, 3)
                         - np.asarray(init["colors"]).reshape(-1, 3), -4, 4)
    return t


def scene_dataset(ctx, seed: int, n_gt: int, n_seed: int, views: int, width: int, height: int,
                  iterations: int):
    sc = mesh_scene(seed, n_gt, n_seed, views, width, height)
    images = render_targets(ctx, sc["gt"], sc["cams"])
    init = U.points_knn_init(ctx, sc["xyz"], sc["rgb"])
    conv = converge(ctx, init, sc["cams"], images, sc["extent"], iterations)
    feats = U.points_features(ctx, sc["xyz"], sc["rgb"], sc["cams"])
    return feats, targets_from(init, conv)


def train_network(ctx, net: U.InitNet, feats: np.ndarray, targets: np.ndarray, steps: int,
                  batch: int = 8192, lr: float = 1e-3, seed: int = 0, log=None) -> list:
    """Adam on random mini-batches; features and targets stay resident on the device."""
    g = np.random.default_rng(seed)
    n = len(feats)
    losses = []
    for it in range(steps):
        idx = g.integers(0, n, size=min(batch, n))
        f = U.DeviceArray.from_numpy(ctx, np.ascontiguousarray(feats[idx], np.float32))
        t = U.DeviceArray.from_numpy(ctx, np.ascontiguousarray(targets[idx], np.float32))
        rate = lr * 0.5 * (1 + math.cos(math.pi * it / max(steps, 1)))
        losses.append(net.train_step(f, t, OUTPUT_WEIGHTS, rate))
        if log and (it % 200 == 0 or it == steps - 1):
            log(f"  step {it} loss {losses[-1]:.4f}")
    return losses


def save_weights(net: U.InitNet, path: str = WEIGHTS, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=net.params(), meta=json.dumps(meta))


def load_weights(ctx, path: str = WEIGHTS) -> "U.InitNet | None":
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        p = z["params"]
    net = U.InitNet(ctx)
    net.set_params(p)
    return net


# ---------------------------------------------------------------------------- evaluation
def time_to_l1(ctx, init: dict, cams, images, extent, opt, target_l1: float, max_steps: int,
               check_every: int = 25, init_seconds: float = 0.0):
    """Wall time (training + init, evaluation passes excluded) until the mean L1 over all views
    is at or below target_l1. Returns (seconds or None, steps, final L1)."""
    t = U.Trainer(ctx, init, cams, images, extent, opt)
    spent, steps = init_seconds, 0
    l1 = t.evaluate_l1()
    while steps < max_steps:
        if l1 <= target_l1:
            return spent, steps, l1
        t0 = time.perf_counter()
        for _ in range(check_every):
            if t.done():
                break
            t.step()
            steps += 1
        t.synchronize()
        spent += time.perf_counter() - t0
        l1 = t.evaluate_l1()
        if t.done():
            break
    return (spent if l1 <= target_l1 else None), steps, l1


def batched_options(plain_iterations: int, views_per_step: int, max_gaussians: int,
                    views_factor: float = 1.5):
    """The plain schedule (TrainOptions::forIterations) counted in views: B views per step,
    densify steps divided by B, sqrt(B) rates and beta^B decay (batch_scaling), the means rate
    decayed over views_factor x the plain trainer's views (also the most it runs), and the
    opacity resets at the plain schedule's fractions of that horizon (every views_factor x the
    plain interval, in views).
    Swept on two held-out high-preset scenes at B = 8: rates x B^0.75 or x B never reached the
    plain trainer's final L1, x B^0.4 or x B^0.3 took 4-20% longer than x sqrt(B); a decay over
    1.25 or 1.5 x the plain views reached it 5-8% sooner than over 2 x, over 1 x not at all.
    On four held-out scenes (kNN and learned starts): resets every 1.5 x the plain interval
    reached it 6-18% sooner than at the plain interval (whose second reset falls shortly before
    the batched run reaches the target), and as soon as with no reset at all; stretching the
    densify schedule by 1.5 x as well was up to 12% slower than even the plain reset interval."""
    B = views_per_step
    base = U.options_for_iterations(plain_iterations)
    o = U.options_for_iterations(max(1, int(plain_iterations * views_factor) // B))
    o.refine_start = max(1, base.refine_start // B)
    o.refine_stop = max(1, base.refine_stop // B)
    o.refine_interval = max(1, base.refine_interval // B)
    o.opacity_reset_interval = (max(1, int(base.opacity_reset_interval * views_factor) // B)
                                if base.opacity_reset_interval else 0)
    o.max_gaussians = max_gaussians
    o.views_per_step = B
    o.use_graph = int(B > 1)
    o.batch_scaling = int(B > 1)
    return o


VARIANTS = (("learned_batched", True, 8), ("knn_batched", False, 8), ("learned_plain", True, 1))


def evaluate_scene(ctx, net, seed: int, n_gt: int, n_seed: int, views: int, width: int,
                   height: int, plain_iterations: int, variants=VARIANTS) -> dict:
    """Plain trainer (kNN init, one view per step) for plain_iterations -> target L1 (its final
    mean L1 over all views); then, per variant (init, views per step), the wall time to reach it
    (initialisation included, evaluation passes excluded)."""
    sc = mesh_scene(seed, n_gt, n_seed, views, width, height)
    images = render_targets(ctx, sc["gt"], sc["cams"])
    cams, extent = sc["cams"], sc["extent"]

    def init_of(use_net):
        ctx.synchronize()
        t0 = time.perf_counter()
        m = U.learned_init(ctx, net if use_net else None, sc["xyz"], sc["rgb"], cams)
        return m, time.perf_counter() - t0

    knn, t_knn = init_of(False)
    opt = U.options_for_iterations(plain_iterations)
    opt.max_gaussians = n_gt
    t = U.Trainer(ctx, knn, cams, images, extent, opt)
    l1_start = t.evaluate_l1()
    ctx.synchronize()
    t0 = time.perf_counter()
    while not t.done():
        t.step()
    t.synchronize()
    plain_s = time.perf_counter() - t0 + t_knn
    target = t.evaluate_l1()
    out = dict(seed=seed, n_seed=n_seed, n_gt=n_gt, views=views, width=width, height=height,
               plain_iterations=plain_iterations, plain_seconds=round(plain_s, 3),
               target_l1=target, knn_initial_l1=l1_start)
    learned, t_learned = init_of(True)
    out["learned_initial_l1"] = U.Trainer(ctx, learned, cams, images, extent,
                                          U.options_for_iterations(10)).evaluate_l1()
    for name, use_net, B in variants:
        model, t_init = (learned, t_learned) if use_net else (knn, t_knn)
        o = batched_options(plain_iterations, B, n_gt)
        check = max(1, 100 // B)
        sec, steps, l1 = time_to_l1(ctx, model, cams, images, extent, o, target, o.iterations,
                                    check_every=check, init_seconds=t_init)
        out[name] = dict(views_per_step=B, seconds=None if sec is None else round(sec, 3),
                         steps=steps, views=steps * B, final_l1=l1,
                         fraction_of_plain=None if sec is None else round(sec / plain_s, 4))
    return out


def sweep_batched(ctx, net, seed: int, specs, plain_iterations: int = 7000, n_gt: int = 300_000,
                  n_seed: int = 50_000, views: int = 200, width: int = 640, height: int = 480,
                  log=print) -> dict:
    """Wall time to the plain trainer's final L1 on one held-out scene for batched variants
    spec = (init "knn" | "learned", views per step B, decay horizon views_factor, rate
    exponent e[, reset factor r[, densify factor d]]): batched_options(plain_iterations, B, n_gt,
    views_factor) with the rates multiplied by B^e (B^0.5 is the engine's own scaling), the
    opacity resets every r x the plain interval in views (default views_factor; 0: none) and
    the densify schedule stretched by d (default 1)."""
    sc = mesh_scene(seed, n_gt, n_seed, views, width, height)
    images = render_targets(ctx, sc["gt"], sc["cams"])
    cams, extent = sc["cams"], sc["extent"]
    ctx.synchronize()
    t0 = time.perf_counter()
    knn = U.learned_init(ctx, None, sc["xyz"], sc["rgb"], cams)
    t_knn = time.perf_counter() - t0
    learned = U.learned_init(ctx, net, sc["xyz"], sc["rgb"], cams) if net else None
    opt = U.options_for_iterations(plain_iterations)
    opt.max_gaussians = n_gt
    t = U.Trainer(ctx, knn, cams, images, extent, opt)
    ctx.synchronize()
    t0 = time.perf_counter()
    while not t.done():
        t.step()
    t.synchronize()
    plain_s = time.perf_counter() - t0 + t_knn
    target = t.evaluate_l1()
    del t
    out = dict(seed=seed, plain_seconds=round(plain_s, 3), target_l1=target, runs=[])
    base = U.options_for_iterations(plain_iterations)
    for spec in specs:
        init, B, vf, ex = spec[:4]
        rf = spec[4] if len(spec) > 4 else vf
        df = spec[5] if len(spec) > 5 else 1.0
        o = batched_options(plain_iterations, B, n_gt, views_factor=vf)
        m = float(B) ** ex / float(B) ** 0.5
        for f in ("means", "scales", "quats", "opacities", "colors"):
            setattr(o.lr, f, getattr(o.lr, f) * m)
        o.opacity_reset_interval = (max(1, int(base.opacity_reset_interval * rf) // B)
                                    if base.opacity_reset_interval and rf > 0 else 0)
        o.refine_start = max(1, int(base.refine_start * df) // B)
        o.refine_stop = max(1, int(base.refine_stop * df) // B)
        o.refine_interval = max(1, int(base.refine_interval * df) // B)
        model = learned if init == "learned" and learned is not None else knn
        sec, steps, l1 = time_to_l1(ctx, model, cams, images, extent, o, target, o.iterations,
                                    check_every=max(1, 100 // B), init_seconds=t_knn)
        rec = dict(init=init, views_per_step=B, views_factor=vf, rate_exponent=ex, reset_factor=rf,
                   densify_factor=df,
                   seconds=None if sec is None else round(sec, 3), steps=steps, views=steps * B,
                   fraction_of_plain=None if sec is None else round(sec / plain_s, 4), final_l1=l1)
        out["runs"].append(rec)
        log(json.dumps(rec))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--train-scenes", type=int, default=48)
    ap.add_argument("--eval-scenes", type=int, default=4)
    ap.add_argument("--train-seeds", type=int, nargs=2, default=(20000, 40000))
    ap.add_argument("--converge-iterations", type=int, default=1500)
    ap.add_argument("--net-steps", type=int, default=3000)
    ap.add_argument("--eval-gt", type=int, default=300000)
    ap.add_argument("--eval-seeds", type=int, default=50000)
    ap.add_argument("--eval-views", type=int, default=200)
    ap.add_argument("--eval-size", type=int, nargs=2, default=(640, 480))
    ap.add_argument("--plain-iterations", type=int, default=7000)
    ap.add_argument("--skip-train", action="store_true", help="evaluate the saved weights")
    ap.add_argument("--out", default=None, help="write the evaluation JSON here")
    ap.add_argument("--weights", default=WEIGHTS, help="network weights file (.npz)")
    ap.add_argument("--sweep", default=None, metavar="SPECS",
                    help="only sweep batched options on held-out scenes: 'init,B,views_factor,"
                         "rate_exponent[,reset_factor[,densify_factor]];...' (init knn or learned)")
    ap.add_argument("--sweep-seeds", type=int, nargs="+", default=[100_000])
    a = ap.parse_args(argv)
    ctx = U.Context(0)
    log = lambda m: print(m, flush=True)
    if a.sweep:
        specs = []
        for item in a.sweep.split(";"):
            f = item.split(",")
            specs.append((f[0], int(f[1]), *[float(x) for x in f[2:]]))
        net = load_weights(ctx, a.weights)
        res = [sweep_batched(ctx, net, sd, specs, a.plain_iterations, a.eval_gt, a.eval_seeds,
                             a.eval_views, a.eval_size[0], a.eval_size[1], log=log)
               for sd in a.sweep_seeds]
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f, indent=1)
        return 0
    if not a.skip_train:
        g = np.random.default_rng(7)
        feats, targs = [], []
        t0 = time.perf_counter()
        for k in range(a.train_scenes):
            ns = int(g.integers(a.train_seeds[0], a.train_seeds[1] + 1))
            f, t = scene_dataset(ctx, k, 2 * ns, ns, 24, 320, 240, a.converge_iterations)
            feats.append(f)
            targs.append(t)
            log(f"scene {k}: {ns} seeds, converged ({time.perf_counter() - t0:.1f} s)")
        feats, targs = np.concatenate(feats), np.concatenate(targs)
        net = U.InitNet(ctx, seed=1)
        t0 = time.perf_counter()
        losses = train_network(ctx, net, feats, targs, a.net_steps, log=log)
        log(f"trained {a.net_steps} steps on {len(feats)} points in {time.perf_counter() - t0:.1f} s; "
            f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
        save_weights(net, a.weights, train_scenes=a.train_scenes, points=int(len(feats)),
                     net_steps=a.net_steps, first_loss=float(losses[0]), last_loss=float(losses[-1]))
    else:
        net = load_weights(ctx, a.weights)
        assert net is not None, f"no weights at {a.weights}"
    results = []
    for k in range(a.eval_scenes):
        r = evaluate_scene(ctx, net, 100_000 + k, a.eval_gt, a.eval_seeds, a.eval_views,
                           a.eval_size[0], a.eval_size[1], a.plain_iterations)
        log(json.dumps(r))
        results.append(r)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(results, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
