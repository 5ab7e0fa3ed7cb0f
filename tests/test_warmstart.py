"""Learned warm start: feature gather (pdr_bake_features) and the MLP kernels
(pdr_mlp_infer / pdr_mlp_gradient / pdr_mlp_adam_step) against the oracle.

CPU, pinning the oracle (oracle/mlp.py and orc_bake_features):
* the fp64 gradient equals central finite differences (h = 1e-6);
* bf16 rounding equals torch's float32 -> bfloat16 conversion, ties to even;
* a zero output layer predicts the samples' confidence-weighted mean exactly;
* the first Adam step moves every parameter by lr * g / (|g| + eps);
* features of an orthographic quad seen by 10 views of constant colour:
  the 8 highest confidences in descending order, ties in view order, cos = 1,
  and with at most 8 views the base colour equals the weighted-average bake.
GPU, through the C ABI:
* features bit-exact against the oracle (36 views: two launches carry the lists);
* inference within 2e-3 absolute of the oracle (bf16 operands; the oracle rounds
  at the same points, so the residue is float32 vs float64 accumulation plus
  rare one-ulp bf16 rounding flips);
* loss within 1e-5 relative, gradient blocks within 1e-4 relative L2 and
  elementwise within 1e-3 of the block's largest value;
* an Adam step within 1e-6 of the oracle's;
* training on GPU lowers the loss of a synthetic task.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import mlp as om
from photara_drender_amd import scenes

ORTHO = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, -0.1, 0], [0, 0, 0, 1]], np.float32)


def quad():
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return pos, nrm, uv, idx


def random_features(n, rng, empty=0.1):
    x = np.zeros((n, 40), np.float32)
    for t in range(n):
        k = 0 if rng.uniform() < empty else rng.integers(1, 9)
        conf = np.sort(rng.uniform(0.05, 1.0, k))[::-1]
        for s in range(k):
            x[t, 5 * s:5 * s + 3] = rng.uniform(0, 1, 3)
            x[t, 5 * s + 3] = conf[s]
            x[t, 5 * s + 4] = rng.uniform(0.1, 1.0)
    return x


def random_params(seed):
    rng = np.random.default_rng(seed)
    th = om.init_params(seed).astype(np.float64)
    p = om.unpack(th)
    p["b1"] = rng.normal(0, 0.1, om.HID)
    p["b2"] = rng.normal(0, 0.1, om.HID)
    p["W3"] = rng.normal(0, 0.1, (om.HID, om.OUT))
    p["b3"] = rng.normal(0, 0.05, om.OUT)
    return om.pack(p).astype(np.float32)


def test_oracle_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    th = random_params(1).astype(np.float64)
    x = random_features(24, rng)
    t = rng.uniform(0, 1, (24, 3))
    v = (rng.uniform(size=24) > 0.3).astype(np.float64)
    _, g, _ = om.loss_and_grad(th, x, t, v, bf=False)
    h = 1e-6
    for i in rng.choice(om.PARAMS, 60, replace=False):
        e = np.zeros(om.PARAMS)
        e[i] = h
        fd = (om.loss_and_grad(th + e, x, t, v, bf=False)[0] - om.loss_and_grad(th - e, x, t, v, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-8 + 1e-6 * abs(g[i]), (i, fd, g[i])


def test_oracle_bf16_rounding_is_round_to_nearest_even():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    a = np.concatenate([rng.normal(0, 3, 4000), rng.normal(0, 1e-3, 1000),
                        1.0 + np.arange(8) * 2.0 ** -8]).astype(np.float32)  # exact ties at odd multiples
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(om.bf16(a).astype(np.float32), ref)
    assert om.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0 and om.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def test_zero_output_layer_predicts_the_base_colour():
    rng = np.random.default_rng(3)
    x = random_features(50, rng)
    y = om.predict(om.init_params(4), x)
    f = x.reshape(-1, 8, 5).astype(np.float64)
    w = f[:, :, 3]
    s = w.sum(1)
    base = (f[:, :, :3] * w[:, :, None]).sum(1) / np.where(s > 0, s, 1)[:, None]
    np.testing.assert_allclose(y, np.where(s[:, None] > 0, base, 0.0), rtol=1e-12, atol=0)


def test_oracle_first_adam_step_closed_form():
    rng = np.random.default_rng(5)
    th = rng.normal(size=100)
    g = rng.normal(size=100)
    new, m, v = om.adam(th, g, np.zeros(100), np.zeros(100), 1, lr=1e-2)
    np.testing.assert_allclose(th - new, 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def ortho_views(colors, weights):
    imgs = [np.full((16, 16, 3), c, np.float32) for c in colors]
    return imgs, np.stack([ORTHO] * len(colors)), [[0, 0, 1000.0]] * len(colors), weights


def test_features_keep_the_eight_most_confident_samples():
    pos, nrm, uv, idx = quad()
    colors = np.linspace(0.05, 0.95, 10)
    weights = [0.5, 2.0, 1.0, 2.0, 0.25, 3.0, 1.5, 0.75, 0.1, 1.0]
    imgs, mats, cams, w = ortho_views(colors, weights)
    f = oracle.bake_features(pos, nrm, uv, idx, imgs, mats, cams, resolution=(16, 16), weights=w)
    order = sorted(range(10), key=lambda k: (-weights[k], k))[:8]  # descending, ties in view order
    slots = f.reshape(16, 16, 8, 5)
    for s, k in enumerate(order):
        np.testing.assert_allclose(slots[..., s, 0], colors[k], atol=1e-6)
        np.testing.assert_allclose(slots[..., s, 3], weights[k], rtol=1e-6)
    np.testing.assert_allclose(slots[..., 4], 1.0, atol=1e-6)  # head-on orthographic view: cos = 1


def test_features_base_colour_equals_weighted_average_bake():
    sc = scenes.atlas_scene(nx=20, ny=10, charts_x=2, charts_y=2, atlas_size=48, gutter=2, seed=6)
    mats = scenes.atlas_views(sc, 6, (40, 56), seed=6)
    rng = np.random.default_rng(6)
    imgs = [rng.uniform(0, 1, (40, 56, 3)).astype(np.float32) for _ in range(6)]
    cams = np.stack([scenes.camera_center(m) for m in mats])
    nrm = scenes.vertex_normals(sc.positions, sc.indices)
    f = oracle.bake_features(sc.positions, nrm, sc.uv, sc.indices, imgs, mats, cams, resolution=(48, 48))
    color, conf, *_ = oracle.bake(sc.positions, nrm, sc.uv, sc.indices, imgs, mats, cams, resolution=(48, 48))
    has = conf > 0
    assert has.sum() > 500
    base = om.base_color(f.reshape(-1, 40)).reshape(48, 48, 3)
    np.testing.assert_allclose(base[has], color[..., :3][has], atol=2e-6)
    np.testing.assert_allclose(f[..., 3::5].sum(-1)[has], conf[has], rtol=2e-6)
    assert (f[~has] == 0).all()


# ---------------------------------------------------------------------- GPU


def _ctx():
    import photara_drender_amd._native as nat

    return nat.Context(0, flags=0)


@pytest.mark.gpu
def test_gpu_features_match_oracle():
    import torch

    from photara_drender_amd.bake import bake_features

    sc = scenes.atlas_scene(nx=40, ny=20, charts_x=4, charts_y=2, atlas_size=64, gutter=2, seed=7)
    views = 36  # more than one 32-view launch
    mats = scenes.atlas_views(sc, views, (24, 32), seed=7)
    rng = np.random.default_rng(7)
    imgs = [rng.uniform(0, 1, (24, 32, 3 + (k % 2))).astype(np.float32) for k in range(views)]
    masks = [rng.uniform(0, 1, (24, 32)).astype(np.float32) if k % 3 == 0 else None for k in range(views)]
    cams = np.stack([scenes.camera_center(m) for m in mats])
    nrm = scenes.vertex_normals(sc.positions, sc.indices)
    w = rng.uniform(0.5, 2.0, views).tolist()
    ref = oracle.bake_features(sc.positions, nrm, sc.uv, sc.indices, imgs, mats, cams, masks, resolution=(64, 64),
                               weights=w, mask_floor=0.1)
    got = bake_features(_ctx(), torch.tensor(sc.positions, device="cuda"), nrm, sc.uv, sc.indices, imgs, mats, cams,
                        masks, resolution=(64, 64), weights=w, mask_floor=0.1).cpu().numpy()
    assert (ref[..., 3] > 0).sum() > 1000 and (ref[..., 38] > 0).sum() > 100  # full lists occur
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def _mlp(params):
    from photara_drender_amd.warmstart import WarmStartMLP

    return WarmStartMLP(_ctx(), params)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 37, 1000])
def test_gpu_mlp_inference_matches_oracle(n):
    import torch

    rng = np.random.default_rng(n)
    x = random_features(n, rng)
    th = random_params(n)
    got = _mlp(th).predict(torch.tensor(x, device="cuda"), clamp=None).cpu().numpy()
    ref = om.predict(th, x)
    np.testing.assert_allclose(got, ref, atol=2e-3, rtol=0)
    assert np.abs(got - ref).mean() < 2e-5
    assert (got[x[:, 3::5].sum(1) == 0] == 0).all()


@pytest.mark.gpu
@pytest.mark.parametrize("n,with_valid", [(1000, True), (77, False), (4133, True)])
def test_gpu_mlp_gradient_matches_oracle(n, with_valid):
    import torch

    rng = np.random.default_rng(n)
    x = random_features(n, rng)
    th = random_params(n + 1)
    target = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    valid = (rng.uniform(size=n) > 0.25).astype(np.float32) if with_valid else None
    m = _mlp(th)
    g, loss, count = m.gradient(torch.tensor(x, device="cuda"), torch.tensor(target, device="cuda"),
                                None if valid is None else torch.tensor(valid, device="cuda"))
    g = g.cpu().numpy().astype(np.float64)
    rloss, rg, _ = om.loss_and_grad(th, x, target, valid)
    assert count == (n if valid is None else valid.sum())
    assert loss == pytest.approx(rloss, rel=1e-5)
    o = 0
    for name, shape in om.SHAPES:
        k = int(np.prod(shape))
        a, b = g[o:o + k], rg[o:o + k]
        o += k
        assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b) + 1e-9, name
        assert np.abs(a - b).max() <= 1e-3 * np.abs(b).max() + 1e-9, name


@pytest.mark.gpu
def test_gpu_mlp_empty_batch_gives_zero_gradient():
    import torch

    m = _mlp(random_params(9))
    g, loss, count = m.gradient(torch.zeros((0, 40), device="cuda"), torch.zeros((0, 3), device="cuda"))
    assert count == 0 and loss == 0 and float(g.abs().max()) == 0


@pytest.mark.gpu
def test_gpu_adam_step_matches_oracle():
    import torch

    rng = np.random.default_rng(11)
    n = 2000
    x = random_features(n, rng)
    th = random_params(12)
    target = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    m = _mlp(th)
    m1 = rng.normal(0, 1e-3, om.PARAMS).astype(np.float32)
    m2 = rng.uniform(1e-6, 1e-5, om.PARAMS).astype(np.float32)
    m.m.copy_(torch.tensor(m1))
    m.v.copy_(torch.tensor(m2))
    m.step = 4
    m.train_step(torch.tensor(x, device="cuda"), torch.tensor(target, device="cuda"), None, lr=1e-3)
    _, rg, _ = om.loss_and_grad(th, x, target)
    new, _, _ = om.adam(th.astype(np.float64), rg, m1.astype(np.float64), m2.astype(np.float64), 5, lr=1e-3)
    np.testing.assert_allclose(m.params.cpu().numpy(), new, atol=1e-6, rtol=0)


@pytest.mark.gpu
def test_gpu_training_lowers_the_loss():
    import torch

    rng = np.random.default_rng(13)
    n = 20000
    x = random_features(n, rng, empty=0.0)
    # target: the most confident sample's colour, darkened by its cosine
    target = (x[:, 0:3] * (0.5 + 0.5 * x[:, 4:5])).astype(np.float32)
    m = _mlp(om.init_params(0))
    xt, tt = torch.tensor(x, device="cuda"), torch.tensor(target, device="cuda")
    _, before, _ = m.gradient(xt, tt)
    for step in range(300):
        sel = torch.randint(0, n, (4096,), device="cuda")
        m.train_step(xt[sel], tt[sel], None, lr=3e-3)
    _, after, _ = m.gradient(xt, tt)
    assert after < 0.25 * before, (before, after)


@pytest.mark.gpu
def test_gpu_trained_network_gain_on_a_held_out_scene():
    """What the shipped network gains over the method without it, on a scene outside
    its training seeds: its own objective (per-texel squared error to the plain
    method's refined atlas) against the untrained network (zero output layer = the
    samples' confidence-weighted mean), and the warm-started refinement reaching
    the plain method's PSNR - 0.1 dB, next to the same schedule from the plain
    initial atlas."""
    from photara_drender_amd import _native, warmstart, workloads

    ctx = _native.Context(0, flags=0)
    prob = workloads.refine_problem(ctx, seed=4321, **warmstart.TRAIN_SCENE)
    init, _, _ = workloads.bake_initial_atlas(ctx, prob)
    refined = warmstart.refine_problem_atlas(ctx, prob, init).color
    feats = warmstart.problem_features(ctx, prob)
    has = feats[..., 3::5].sum(-1) > 0
    trained = warmstart.WarmStartMLP.load(ctx)
    untrained = warmstart.WarmStartMLP(ctx, warmstart.init_params(0))
    _, loss_trained, _ = trained.gradient(feats[has], refined[has])
    _, loss_base, _ = untrained.gradient(feats[has], refined[has])
    ev = warmstart.evaluate(ctx, prob, trained, warm=dict(learning_rate=3e-3))
    print(f"held-out loss: trained {loss_trained:.4e}, untrained {loss_base:.4e}; steps to target: plain "
          f"{ev['plain']['steps_to_target']}, warm {ev['warm_start']['steps_to_target']}, ablation "
          f"{ev['ablation_plain_init_warm_options']['steps_to_target']}; init PSNR plain "
          f"{ev['plain']['init_psnr_db']:.2f} warm {ev['warm_start']['init_psnr_db']:.2f}")
    assert loss_trained < loss_base
    warm_steps = ev["warm_start"]["steps_to_target"]
    ablation_steps = ev["ablation_plain_init_warm_options"]["steps_to_target"]
    assert warm_steps is not None
    assert ablation_steps is None or warm_steps <= ablation_steps
