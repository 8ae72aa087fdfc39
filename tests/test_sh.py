"""Spherical-harmonic colour of degree 1..3 (a build-side extension: the reference trains the
DC term only, GaussianModel.h:21, so there is no reference output to compare against).

Pinned by mathematics on the CPU oracle: the 16 basis functions are orthonormal on the
sphere (a wrong constant, sign or dropped term breaks it); degree 3 with zero coefficients
reduces bit-exactly to the reference's DC colour; analytic gradients of the coefficients and of
the means (through the view direction) match central differences of an FP64 forward. The GPU
path is then compared with the oracle: forward bit-exact, gradients within 1e-4.
"""
import math

import numpy as np
import pytest

import oracle as O
from scenes import config0_scene, fd_camera, gaussians, ref_style_model
from univeral_3d_reconstruction_system_amd import synthetic as S

FIELDS = [f for f, _ in O.FIELDS]


def test_sh_basis_orthonormal():
    # Gauss-Legendre in cos(theta) x uniform phi integrates degree <= 6 polynomials exactly.
    nt, nphi = 16, 32
    ct, wt = np.polynomial.legendre.leggauss(nt)
    phis = 2 * math.pi * np.arange(nphi) / nphi
    basis, weights = [], []
    for c, w in zip(ct, wt):
        s = math.sqrt(1 - c * c)
        for p in phis:
            b = O.sh_basis(s * math.cos(p), s * math.sin(p), c)
            basis.append(np.concatenate([[0.28209479177387814], b]))
            weights.append(w * 2 * math.pi / nphi)
    B = np.array(basis)
    G = (B * np.array(weights)[:, None]).T @ B
    assert np.abs(G - np.eye(16)).max() < 2e-6


def test_degree3_with_zero_coefficients_is_the_reference_colour():
    g, cams = config0_scene(3000, 4, 96, 72)
    st = O.make_settings()
    a = O.Raster().forward(gaussians(g), cams[1], st)
    b = O.Raster().forward(gaussians(g), cams[1], st, sh_rest=np.zeros(3000 * 45, np.float32), sh_degree=3)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("degree", [1, 3])
def test_sh_gradients_match_fp64_finite_differences(degree):
    g = ref_style_model(6, 13, 0.6, 3.0, 0.08, 0.3)
    rest = (S.rng(14).normal(0, 0.3, 6 * 45)).astype(np.float32)
    cam = fd_camera()
    st = O.make_settings(tile_size=64, min_alpha=0.0, min_transmittance=0.0, background=(0.1, 0.2, 0.3))
    w = np.random.default_rng(3).uniform(-1, 1, cam["width"] * cam["height"] * 3).astype(np.float32)
    r = O.Raster()
    r.forward(gaussians(g), cam, st, sh_rest=rest, sh_degree=degree)
    grads, grest = r.backward_sh(gaussians(g), w, rest, degree)
    g64 = {k: v.astype(np.float64) for k, v in g.items()}
    rest64 = rest.astype(np.float64)
    used = [i * 45 + k for i in range(6) for k in range({1: 9, 3: 45}[degree])]

    def loss(gg, rr):
        return float(np.dot(w.astype(np.float64), O.render_f64_sh(gg, rr, degree, cam, st)))
    eps = 1e-5
    num = np.array([(loss(g64, rest64 + eps * (np.arange(270) == k)) -
                     loss(g64, rest64 - eps * (np.arange(270) == k))) / (2 * eps) for k in used])
    assert np.linalg.norm(grest[used] - num) / np.linalg.norm(num) < 2e-3
    assert np.all(np.delete(grest, used) == 0)
    nm = np.zeros(18)
    for k in range(18):
        gp = {kk: vv.copy() for kk, vv in g64.items()}
        gm = {kk: vv.copy() for kk, vv in g64.items()}
        gp["means"][k] += eps
        gm["means"][k] -= eps
        nm[k] = (loss(gp, rest64) - loss(gm, rest64)) / (2 * eps)
    assert np.linalg.norm(grads.means - nm) / np.linalg.norm(nm) < 2e-3


# ------------------------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def ctx():
    from univeral_3d_reconstruction_system_amd import splat as U
    c = U.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_gpu_sh_forward_bit_exact_backward_close(ctx, degree):
    from univeral_3d_reconstruction_system_amd import splat as U
    g, cams = config0_scene(8000, 6, 160, 120)
    n = 8000
    rest = S.rng(7).normal(0, 0.05, n * 45).astype(np.float32)
    cam, st = cams[5], O.make_settings()
    ro = O.Raster()
    ref = ro.forward(gaussians(g), cam, st, sh_rest=rest, sh_degree=degree)
    dg = U.DeviceGaussians.from_host(ctx, g)
    drest = U.DeviceArray.from_numpy(ctx, rest)
    r = U.Rasterizer(ctx)
    img = r.forward(dg, cam, U.make_settings(), sh_rest=drest, sh_degree=degree).numpy()
    assert np.array_equal(img, ref)
    dl = np.random.default_rng(2).uniform(-1e-3, 1e-3, img.size).astype(np.float32)
    go, gro = ro.backward_sh(gaussians(g), dl, rest, degree)
    grads = U.DeviceGaussians(ctx, n)
    grest = U.DeviceArray.from_numpy(ctx, np.zeros(n * 45, np.float32))
    r.backward(dg, U.DeviceArray.from_numpy(ctx, dl), grads, sh_rest=drest, sh_degree=degree, grad_sh_rest=grest)
    gh = grads.to_host()
    rel = lambda a, b: np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-30)
    for f in FIELDS:
        assert rel(gh[f], getattr(go, f)) < 1e-4, f
    assert rel(grest.numpy(), gro) < 1e-4


@pytest.mark.gpu
def test_gpu_trainer_with_sh3_learns_view_dependence(ctx):
    from univeral_3d_reconstruction_system_amd import splat as U
    truth = ref_style_model(300, 21, 1.0, 0.0, 0.03, 0.12)
    truth["opacities"][:] = np.float32(math.log(0.8 / 0.2))
    rest_t = S.rng(22).normal(0, 0.15, 300 * 45).astype(np.float32)
    ro = O.Raster()
    cams, imgs = [], []
    for k in range(16):
        cam = S.orbit_camera(96, 72, 80.0, 6.2831853 * k / 16.0, 4.0, -1.0 + 0.15 * (k % 4))
        cams.append(cam)
        imgs.append(S.quantize(ro.forward(gaussians(truth), cam, O.make_settings(), sh_rest=rest_t, sh_degree=3)))
    init = {k: v.copy() for k, v in truth.items()}
    psnr = {}
    for deg in (0, 3):
        opt = U.options_for_iterations(800)
        opt.refine_start = 10**6  # no densification: same Gaussians, only colours differ
        opt.sh_degree = deg
        t = U.Trainer(ctx, init, cams, imgs, 4.4, opt)
        while not t.done():
            t.step()
        psnr[deg] = t.evaluate_psnr()
        if deg == 3:
            assert np.abs(t.sh_rest()).max() > 0
    assert psnr[3] > psnr[0] + 1.0, psnr
