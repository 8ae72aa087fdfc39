"""Batched engine (B views per step, scene sets, whole-step hipGraph) against the plain path.

The batched engine has no reference counterpart (the reference trains one view per step,
Trainer.cpp:64). Its contract is the plain path's mathematics:
  * one step's gradient is the mean of the B per-view gradients of the plain single-view path
    (which matches the oracle, test_gpu_parity.py) -- within 1e-4 relative (L2 per parameter
    group): the compositing backward's float atomics and the order of the per-view sums are
    the only differences;
  * the first step's loss only depends on the forward pass and the loss, whose kernels are the
    plain path's: bit-identical;
  * graph replay and eager launches run the same kernels: same gradients to the same bar;
  * a scene inside a scene set trains as it does alone.
The gradient is read from Adam's first moment after one step with all learning rates at 1e-30
(m = (1 - 0.9f) * g; the parameters do not move).
"""
import numpy as np
import pytest

import oracle as O
from scenes import config0_scene, gaussians
from univeral_3d_reconstruction_system_amd import splat as U
from univeral_3d_reconstruction_system_amd import synthetic as S

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-4
ONE_MINUS_B1 = np.float32(1.0) - np.float32(0.9)


@pytest.fixture(scope="module")
def ctx():
    c = U.Context(0)
    yield c
    c.close()


def _scene(n, views, width, height, seed=1):
    truth, cams = config0_scene(n, seed, width, height, views)
    ro, st = O.Raster(), O.make_settings()
    images = [S.quantize(ro.forward(gaussians(truth), c, st)) for c in cams]
    init, _ = config0_scene(n, seed + 1, width, height, views)
    return init, cams, images


def _frozen_options(views, graph, iterations=100):
    opt = U.options_for_iterations(iterations)
    opt.lr = U.learning_rates(1e-30, 1e-30, 1e-30, 1e-30, 1e-30)
    opt.means_lr_final = 1e-30
    opt.refine_start = 10 ** 6
    opt.opacity_reset_interval = 0
    opt.views_per_step = views
    opt.use_graph = int(graph)
    return opt


def _plain_mean_gradient(ctx, init, cams, images):
    """Mean over the views of the plain single-view gradient (raster fwd, loss, raster bwd)."""
    n = len(init["opacities"])
    dg = U.DeviceGaussians.from_host(ctx, init)
    acc = {f: np.zeros(n * d, np.float64) for f, d in U.FIELDS}
    losses = []
    for cam, img in zip(cams, images):
        r = U.Rasterizer(ctx)
        out = r.forward(dg, cam)
        tgt = U.DeviceArray.from_numpy(ctx, img.reshape(-1).astype(np.float32) * (np.float32(1) / np.float32(255)))
        grad = U.DeviceArray(ctx, (cam["height"] * cam["width"] * 3,))
        losses.append(U.photometric_loss(ctx, out, tgt, cam["width"], cam["height"], 0.2, grad)[0])
        gr = U.DeviceGaussians(ctx, n)
        r.backward(dg, grad, gr)
        h = gr.to_host()
        for f, _ in U.FIELDS:
            acc[f] += h[f]
    return {f: acc[f] / len(cams) for f in acc}, losses


def _assert_grad_close(got, want, label):
    for f, _ in U.FIELDS:
        g, w = np.asarray(got[f], np.float64), want[f]
        err = np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-30)
        assert err < GRAD_RTOL, (label, f, err)


@pytest.mark.parametrize("views,graph", [(1, True), (4, False), (4, True)])
def test_batched_gradient_is_mean_of_plain_views(ctx, views, graph):
    init, cams, images = _scene(3000, views, 160, 120)
    t = U.Trainer(ctx, init, cams, images, S.scene_extent(cams), _frozen_options(views, graph))
    t.step()
    m, _ = t.moments()
    got = {f: m[f].astype(np.float64) / np.float64(ONE_MINUS_B1) for f in m}
    want, losses = _plain_mean_gradient(ctx, init, cams, images)
    _assert_grad_close(got, want, (views, graph))
    # One epoch = all views in one step, so the step's loss is the mean of the views' losses.
    assert np.isclose(t.stats()["loss"], np.mean(np.float32(losses)), rtol=1e-6, atol=0)
    c = t.counters()
    assert c["steps"] == 1 and c["views"] == views and c["overflow_steps"] == 0
    # The parameters did not move.
    mod = t.model()
    for f, _ in U.FIELDS:
        np.testing.assert_allclose(mod[f], np.asarray(init[f], np.float32).reshape(-1), rtol=0, atol=1e-6)


def test_single_view_batched_first_loss_bit_identical(ctx):
    """B = 1 through the engine: same view order and the plain path's forward and loss."""
    init, cams, images = _scene(3000, 6, 160, 120)
    opt_plain = _frozen_options(1, False)
    opt_plain.views_per_step = 0
    plain = U.Trainer(ctx, init, cams, images, 3.0, opt_plain)
    batched = U.Trainer(ctx, init, cams, images, 3.0, _frozen_options(1, True))
    for _ in range(3):
        plain.step()
        batched.step()
        assert plain.stats()["loss"] == batched.stats()["loss"]


def test_graph_replays_match_eager_over_steps(ctx):
    """Several replays of one captured graph against eager launches of the same steps."""
    init, cams, images = _scene(4000, 12, 160, 120)
    opt = U.options_for_iterations(400)
    opt.refine_start = 10 ** 6
    opt.views_per_step = 3
    runs = []
    for graph in (False, True):
        opt.use_graph = int(graph)
        t = U.Trainer(ctx, init, cams, images, S.scene_extent(cams), opt)
        for _ in range(8):
            t.step()
        runs.append((t.model(), t.stats()))
    (ma, sa), (mb, sb) = runs
    assert abs(sa["loss"] - sb["loss"]) <= 1e-5 * abs(sa["loss"])
    for f, _ in U.FIELDS:
        np.testing.assert_allclose(ma[f], mb[f], rtol=0, atol=5e-4, err_msg=f)


def test_scene_set_members_train_as_alone(ctx):
    """Two scenes of different size and resolution in one set vs each trained alone."""
    scenes = [_scene(2500, 5, 160, 120, seed=3), _scene(1200, 4, 96, 72, seed=7)]
    views = [2, 1]

    def trainer(k, graph):
        init, cams, images = scenes[k]
        opt = U.options_for_iterations(50)
        opt.refine_start = 10 ** 6
        opt.views_per_step = views[k]
        opt.use_graph = int(graph)
        opt.seed = 40 + k
        return U.Trainer(ctx, init, cams, images, S.scene_extent(cams), opt)

    alone = [trainer(0, True), trainer(1, True)]
    for t in alone:
        t.step()
    s = U.SceneSet(ctx, use_graph=True)
    for k in range(2):
        assert s.add(trainer(k, True))
    assert len(s) == 2 and s.bytes() > 0
    s.step()
    for k in range(2):
        # first step: forward and loss only -> identical
        assert s.member(k).stats()["loss"] == alone[k].stats()["loss"]
    for _ in range(4):
        s.step()
        for t in alone:
            t.step()
    for k in range(2):
        a, b = alone[k].model(), s.member(k).model()
        for f, _ in U.FIELDS:
            np.testing.assert_allclose(a[f], b[f], rtol=0, atol=5e-4, err_msg=(k, f))
        assert s.member(k).iteration() == 5


def test_batched_engine_refuses_tile_sizes_off_multiples_of_16(ctx):
    """The plain trainer takes any tile size; the batched engine's kernels stage 16x16 blocks of
    one tile, so it refuses other sizes with U3D_ERR_UNSUPPORTED instead of failing later."""
    init, cams, images = _scene(500, 2, 64, 48)
    opt = _frozen_options(2, False)
    opt.raster.tile_size = 20
    with pytest.raises(U.U3DError) as e:
        U.Trainer(ctx, init, cams, images, 3.0, opt)
    assert e.value.code == 4  # U3D_ERR_UNSUPPORTED
    plain = _frozen_options(1, False)
    plain.raster.tile_size = 20
    t = U.Trainer(ctx, init, cams, images, 3.0, plain)
    s = U.SceneSet(ctx, use_graph=False)
    with pytest.raises(U.U3DError):
        s.add(t)
    t.step()


def test_scene_set_admission_by_budget(ctx):
    init, cams, images = _scene(2000, 2, 96, 72)
    opt = U.options_for_iterations(10)
    opt.views_per_step = 2
    s = U.SceneSet(ctx, use_graph=False, hbm_budget_bytes=1 << 20)  # 1 MiB: nothing fits
    t = U.Trainer(ctx, init, cams, images, 3.0, opt)
    assert not s.add(t)
    t.step()  # still owned by the caller and usable
    assert t.iteration() == 1


def test_batched_training_with_densification_converges(ctx):
    """B = 4 with the densify / prune schedule and graph capture (re-captured as N changes)."""
    truth = S.random_gaussians(3000, 21)
    cams = S.ring_cameras(16, 128, 96)
    ro, st = O.Raster(), O.make_settings()
    images = [S.quantize(ro.forward(gaussians(truth), c, st)) for c in cams]
    init = S.random_gaussians(1500, 22)
    opt = U.options_for_iterations(600)
    opt.max_gaussians = 6000
    opt.views_per_step = 4
    opt.use_graph = 1
    t = U.Trainer(ctx, init, cams, images, S.scene_extent(cams), opt)
    before = t.evaluate_psnr()
    while not t.done():
        t.step()
    after = t.evaluate_psnr()
    assert t.size() != 1500
    assert after > before + 3.0, (before, after)
    c = t.counters()
    assert c["overflow_steps"] == 0 and c["views"] == 4 * 600


@pytest.mark.parametrize("lds,heavy,per_cu,items,few,radix,pairs", [
    (64, 1, 0, 0, 4, 4, 1), (128, 0, 2, 4, 16, 6, 0), (256, 1, 1, 8, 0, 8, 1), (256, 1, 0, 0, 0, 5, 0),
    (128, 1, 0, 0, 0, 8, 2), (64, 0, 1, 4, 4, 6, 2)])
def test_launch_choices_do_not_change_results(ctx, lds, heavy, per_cu, items, few, radix, pairs):
    """LDS batch, block order, workgroups per CU, radix items per lane, radix digit width and
    how the tile lists are built (difference grids per image in index order with per-tile depth
    sorts, a sort of emitted pairs, or difference grids in the order of a depth sort): same
    forward (first loss bit-identical), same gradient (1e-4)."""
    init, cams, images = _scene(3000, 4, 160, 120)
    t = U.Trainer(ctx, init, cams, images, S.scene_extent(cams), _frozen_options(4, True))
    t.set_tuning(heavy_first=heavy, workgroups_per_cu=per_cu, lds_batch=lds, sort_items=items,
                 atomic_lanes=few, sort_radix=radix, pair_sort=pairs)
    assert t.tuning()["lds_batch"] == lds and t.tuning()["pair_sort"] == pairs
    t.step()
    m, _ = t.moments()
    got = {f: m[f].astype(np.float64) / np.float64(ONE_MINUS_B1) for f in m}
    want, losses = _plain_mean_gradient(ctx, init, cams, images)
    _assert_grad_close(got, want, (lds, heavy, per_cu, items, few))
    ref = U.Trainer(ctx, init, cams, images, S.scene_extent(cams), _frozen_options(4, True))
    ref.step()
    assert t.stats()["loss"] == ref.stats()["loss"]


def test_autotuner_picks_a_configuration_and_training_continues(ctx):
    init, cams, images = _scene(4000, 16, 160, 120)
    opt = U.options_for_iterations(300)
    opt.views_per_step = 4
    opt.use_graph = 1
    t = U.Trainer(ctx, init, cams, images, S.scene_extent(cams), opt)
    chosen = t.autotune(steps_per_candidate=2)
    assert chosen["lds_batch"] in (64, 128, 256) and chosen["tile_size"] == 16
    assert t.tuning() == chosen
    it = t.iteration()
    assert it == (24 + 9) * 3  # 24 + 9 candidates x (1 warm-up + 2 timed) steps
    for _ in range(5):
        t.step()
    assert t.iteration() == it + 5 and t.counters()["overflow_steps"] == 0
    s = U.SceneSet(ctx, use_graph=True)
    opt.seed = 7
    assert s.add(U.Trainer(ctx, init, cams, images, S.scene_extent(cams), opt))
    c2 = s.autotune(steps_per_candidate=1, tune_tile_size=True)
    assert c2["tile_size"] in (16, 32)
