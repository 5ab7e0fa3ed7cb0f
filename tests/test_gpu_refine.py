"""GPU atlas refinement (pdr_refine) against the CPU oracle's restatement of
TextureRefiner::refine (src/texture_refiner.cpp:174-454).

Bar for the reference objective: bit-exact colours, loss and seam-loss
histories.  The GPU adds every texel's gradient terms in the oracle's order
(batch slot, raster pixel, tap; for seam pairs tap then side), sums each
step's loss over the shaders' 64-pixel groups with their tree order and adds
the group sums in group order, and evaluates Adam as atlas_adam_update.hlsl:
27-42 writes it (correctly rounded sqrt and division); the schedule comes from
the same host libm calls.  Graph replay equals the eager loop bit for bit.
"""
import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu


def assert_bits(a, b, what):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    diff = a.view(np.uint32) != b.view(np.uint32)
    assert not diff.any(), f"{what}: {int(diff.sum())} of {a.size} differ, max |d| {np.abs(a - b).max():.3g}"


def quad_scene(n=16):
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    img = np.zeros((n, n, 4), np.float32)
    img[..., 0], img[..., 1], img[..., 3] = 0.75, 0.25, 1.0
    return pos, uv, idx, img


def small_atlas(views=6, size=(72, 96), seed=4):
    sc = scenes.atlas_scene(nx=40, ny=20, charts_x=4, charts_y=2, atlas_size=64, gutter=2, seed=seed)
    mats = scenes.atlas_views(sc, views, size, seed=seed)
    masks = scenes.ellipse_masks(views, size, seed=seed)
    rng = np.random.default_rng(seed)
    images = []
    for k in range(views):
        clip = oracle.transform_positions(sc.positions, mats[k])
        r, _, _ = oracle.rasterize_forward(clip, sc.indices, size, want_deriv=False)
        surf = oracle.interpolate_forward(sc.surface, sc.indices, r)
        gt = scenes.ground_truth(surf.reshape(-1, 2)).reshape(size + (3,))
        img = np.clip(gt + rng.normal(0, 0.02, gt.shape), 0, 1).astype(np.float32)
        images.append(img)
    pairs = scenes.seam_uv_pairs(sc.indices, sc.uv, sc.vertex_remap, sc.face_chart_ids)
    return sc, mats, masks, images, pairs


def test_refine_python_smoke_matches_oracle():
    """tests/python_smoke.py:83-100 through TextureBaker.refine_texture."""
    import photara_drender_amd as pdr

    pos, uv, idx, img = quad_scene()
    init = np.full((16, 16, 3), 0.25, np.float32)
    baker = pdr.TextureBaker()
    color, hist, seam, pre, opt = baker.refine_texture(init, pos, uv, idx, [img], np.eye(4, dtype=np.float32)[None],
                                                       steps=3, batch_size=1, learning_rate=0.05,
                                                       minimum_learning_rate=0.01, seam_polish_steps=0)
    assert color.shape == (16, 16, 3) and hist.shape == (3,) and np.isfinite(hist).all()
    assert hist[-1] < hist[0] and seam.shape == (0,) and pre >= 0 and opt >= 0
    ref_c, ref_h, _ = oracle.refine(init, pos, uv, idx, [img], np.eye(4, dtype=np.float32)[None], steps=3,
                                    batch_size=1, learning_rate=0.05, minimum_learning_rate=0.01,
                                    seam_polish_steps=0)
    assert_bits(color, ref_c, "colour")
    assert_bits(hist, ref_h, "loss history")


def test_refine_atlas_scene_matches_oracle():
    import photara_drender_amd as pdr

    sc, mats, masks, images, pairs = small_atlas()
    init = np.full((64, 64, 3), 0.5, np.float32)
    kw = dict(steps=40, batch_size=2, learning_rate=1e-2, minimum_learning_rate=2.5e-4, photometric_epsilon=1e-3,
              seam_polish_steps=5, seam_learning_rate=1e-3)
    baker = pdr.TextureBaker()
    color, hist, seam, _, _ = baker.refine_texture(init, sc.positions, sc.uv, sc.indices, images, mats,
                                                   visibility_masks=masks, seam_uv_pairs=pairs, **kw)
    ref_c, ref_h, ref_s = oracle.refine(init, sc.positions, sc.uv, sc.indices, images, mats, visibility_masks=masks,
                                        seam_uv_pairs=pairs, **kw)
    assert hist.shape == (40,) and seam.shape == (5,)
    assert_bits(hist, ref_h, "loss history")
    assert_bits(seam, ref_s, "seam loss history")
    assert_bits(color, ref_c, "colour")
    assert hist[-1] < hist[0]


def test_refine_seam_fold_over_many_pairs_matches_oracle():
    """Over 262,144 seam pairs (the pair list repeated): each seam step's loss
    adds more than 4,096 group sums, so the fold stages its row in several
    chunks; the history stays bit-exact."""
    import photara_drender_amd as pdr

    sc, mats, masks, images, pairs = small_atlas(views=2)
    many = np.concatenate([pairs] * (262144 // len(pairs) + 2))
    assert (len(many) + 63) // 64 > 4096
    init = np.full((64, 64, 3), 0.5, np.float32)
    kw = dict(steps=2, batch_size=2, learning_rate=1e-2, seam_polish_steps=3, seam_learning_rate=1e-3)
    baker = pdr.TextureBaker()
    color, hist, seam, _, _ = baker.refine_texture(init, sc.positions, sc.uv, sc.indices, images, mats,
                                                   visibility_masks=masks, seam_uv_pairs=many, **kw)
    ref_c, ref_h, ref_s = oracle.refine(init, sc.positions, sc.uv, sc.indices, images, mats, visibility_masks=masks,
                                        seam_uv_pairs=many, **kw)
    assert seam.shape == (3,) and (seam > 0).all()
    assert_bits(seam, ref_s, "seam loss history")
    assert_bits(hist, ref_h, "loss history")
    assert_bits(color, ref_c, "colour")


def test_refine_graph_matches_eager_and_checkpoints():
    torch = pytest.importorskip("torch")
    from photara_drender_amd import _native
    from photara_drender_amd.refine import refine_atlas

    sc, mats, masks, images, pairs = small_atlas(views=4)
    dev = torch.device("cuda:0")
    ctx = _native.Context(0)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    args = (t(np.full((64, 64, 3), 0.5, np.float32)), t(sc.positions), t(sc.uv), t(sc.indices.view(np.int32)),
            [t(im) for im in images], mats)
    kw = dict(visibility_masks=[t(m) for m in masks], seam_uv_pairs=t(pairs), steps=30, batch_size=2,
              learning_rate=1e-2, seam_polish_steps=4)
    a = refine_atlas(ctx, *args, use_graph=True, checkpoint_every=10, **kw)
    b = refine_atlas(ctx, *args, use_graph=False, **kw)
    assert_bits(a.color.cpu().numpy(), b.color.cpu().numpy(), "colour")
    assert_bits(a.loss_history, b.loss_history, "loss history")
    assert_bits(a.seam_loss_history, b.seam_loss_history, "seam loss history")
    assert a.checkpoints.shape == (3, 64, 64, 3)
    # checkpoints are taken after steps 10, 20, 30 (before the seam polish, which moves the atlas further)
    assert not torch.equal(a.checkpoints[0], a.checkpoints[-1])
    assert not torch.equal(a.checkpoints[-1], a.color)


@pytest.mark.parametrize("tv", [0.0, 1e-3], ids=["reference", "extended"])
def test_refine_stop_after_equals_the_full_runs_checkpoint(tv):
    """stop_after = k runs the first k steps of the same schedule: the atlas equals
    the full run's checkpoint after step k bit for bit, as do the first k losses
    (the rest stay 0), in graph and eager mode, with the TV buffers' parity."""
    torch = pytest.importorskip("torch")
    from photara_drender_amd import _native
    from photara_drender_amd.refine import refine_atlas

    sc, mats, masks, images, pairs = small_atlas(views=4)
    ctx = _native.Context(0)
    init = np.full((64, 64, 3), 0.5, np.float32)
    kw = dict(seam_uv_pairs=pairs, steps=40, batch_size=2, learning_rate=1e-2, seam_polish_steps=0, tv_weight=tv,
              chart_ids=sc.chart_texels if tv else None, joint_seam_weight=1e-2 if tv else 0.0)
    full = refine_atlas(ctx, init, sc.positions, sc.uv, sc.indices, images, mats, masks, checkpoint_every=10, **kw)
    for stop, graph in ((20, True), (30, False), (13, True)):
        part = refine_atlas(ctx, init, sc.positions, sc.uv, sc.indices, images, mats, masks, stop_after=stop,
                            use_graph=graph, **kw)
        assert_bits(part.loss_history[:stop], full.loss_history[:stop], f"losses to step {stop}")
        assert not part.loss_history[stop:].any()
        if stop % 10 == 0:
            assert_bits(part.color, full.checkpoints[stop // 10 - 1], f"atlas after step {stop}")


@pytest.mark.parametrize("size", [(72, 96), (70, 93)])
def test_refine_after_a_bake_resolves_its_shadow_keys(size):
    """After a bake on the same context, pdr_refine resolves the bake's shadow-map
    keys of the same views (same mesh, size, viewport and matrix) instead of
    rasterising them again (no candidate pass runs); the refinement equals one on a
    fresh context bit for bit.  A bake of a moved mesh is not reused.  70 x 93 views
    have padded key tiles (sides not multiples of 4)."""
    torch = pytest.importorskip("torch")
    from photara_drender_amd import _native
    from photara_drender_amd.bake import bake_atlas
    from photara_drender_amd.refine import refine_atlas

    sc, mats, masks, images, pairs = small_atlas(views=4, size=size)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    nrm = scenes.vertex_normals(sc.positions, sc.indices)
    cams = np.stack([scenes.camera_center(m) for m in mats])
    init = np.full((64, 64, 3), 0.5, np.float32)
    kw = dict(seam_uv_pairs=pairs, steps=12, batch_size=2, learning_rate=1e-2, seam_polish_steps=2)
    ref = refine_atlas(_native.Context(0), init, sc.positions, sc.uv, sc.indices, images, mats, masks, **kw)
    for moved in (0.0, 0.01):
        ctx = _native.Context(0)
        bake_atlas(ctx, t(sc.positions + np.float32(moved)), t(nrm), t(sc.uv), t(sc.indices.view(np.int32)),
                   [t(im) for im in images], mats, cams, [t(m) for m in masks], resolution=(64, 64))
        ctx.set_timing(True)
        got = refine_atlas(ctx, init, sc.positions, sc.uv, sc.indices, images, mats, masks, **kw)
        candidates = ctx.kernel_time("raster_candidates")[1] + ctx.kernel_time("raster_tiled")[1]
        ctx.set_timing(False)
        assert_bits(got.color, ref.color, f"colour (mesh moved by {moved})")
        assert_bits(got.loss_history, ref.loss_history, f"loss history (mesh moved by {moved})")
        assert (candidates == 0) == (moved == 0.0), candidates


def test_refine_extended_objective_runs_and_reduces_seams():
    torch = pytest.importorskip("torch")
    from photara_drender_amd import _native
    from photara_drender_amd.refine import refine_atlas

    sc, mats, masks, images, pairs = small_atlas(views=4)
    ctx = _native.Context(0)
    init = np.full((64, 64, 3), 0.5, np.float32)
    base = refine_atlas(ctx, init, sc.positions, sc.uv, sc.indices, images, mats, masks, None, pairs, steps=30,
                        batch_size=2, learning_rate=1e-2, seam_polish_steps=0)
    ext = refine_atlas(ctx, init, sc.positions, sc.uv, sc.indices, images, mats, masks, None, pairs, steps=30,
                       batch_size=2, learning_rate=1e-2, seam_polish_steps=0, tv_weight=1e-3,
                       chart_ids=sc.chart_texels, joint_seam_weight=1e-2)
    assert np.isfinite(ext.color).all() and not np.array_equal(ext.color, base.color)
    eps = 1e-4
    _, l_base = oracle.seam_gradient(base.color, pairs, eps, 1.0 / (3 * pairs.shape[0]))
    _, l_ext = oracle.seam_gradient(ext.color, pairs, eps, 1.0 / (3 * pairs.shape[0]))
    assert l_ext < l_base


def test_refine_validation_messages():
    import photara_drender_amd as pdr

    pos, uv, idx, img = quad_scene()
    baker = pdr.TextureBaker()
    init = np.full((16, 16, 3), 0.25, np.float32)
    eye = np.eye(4, dtype=np.float32)[None]
    with pytest.raises(ValueError, match="texture refinement options are invalid"):
        baker.refine_texture(init, pos, uv, idx, [img], eye, steps=0)
    with pytest.raises(ValueError, match="texture refinement options are invalid"):
        baker.refine_texture(init, pos, uv, idx, [img], eye, learning_rate=1e-3, minimum_learning_rate=1e-2)
    with pytest.raises(ValueError, match="texture refinement mesh arrays have inconsistent shapes"):
        baker.refine_texture(init, pos, uv, np.array([[0, 1, 9]], np.uint32), [img], eye)
    with pytest.raises(ValueError, match="each image must have shape"):
        baker.refine_texture(init, pos, uv, idx, [img[..., :2].copy()], eye)


def test_refine_tv_gradient_matches_autograd():
    """The chart-TV gradient the fused Adam step adds (pdr_testing_tv_gradient) against
    torch autograd (float64) of atlas_total_variation (python/photara_drender/atlas.py:
    84-115), texels with chart id < 0 outside every pair."""
    torch = pytest.importorskip("torch")
    from photara_drender_amd import _native
    from photara_drender_amd.atlas import atlas_total_variation

    rng = np.random.default_rng(21)
    h, w = 29, 37
    tex = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    tex[5:9, 5:9] = 0.5  # flat patch: differences below epsilon
    chart = (np.arange(w)[None, :] // 9 + 4 * (np.arange(h)[:, None] // 10)).astype(np.int32)
    chart[rng.uniform(size=(h, w)) < 0.1] = -1
    eps, weight = 1e-4, 3e-2
    t64 = torch.tensor(tex.astype(np.float64), requires_grad=True)
    c = torch.tensor(chart)
    loss = atlas_total_variation(t64, valid_mask=c >= 0, chart_ids=c, epsilon=eps) * weight
    loss.backward()
    pairs = int(((chart[:, 1:] == chart[:, :-1]) & (chart[:, 1:] >= 0)).sum()
                + ((chart[1:] == chart[:-1]) & (chart[1:] >= 0)).sum())
    dev = torch.device("cuda:0")
    _native.Context(0)
    dt, dc = torch.tensor(tex, device=dev), torch.tensor(chart, device=dev)
    g = torch.empty_like(dt)
    assert _native.lib().pdr_testing_tv_gradient(dt.data_ptr(), w, h, dc.data_ptr(), eps, weight / (pairs * 3),
                                                 g.data_ptr()) == 0
    np.testing.assert_allclose(g.cpu().numpy(), t64.grad.numpy(), rtol=1e-5, atol=1e-9)


def test_refine_graph_matches_eager_with_tv_and_odd_blocks():
    """Graph replay equals the eager loop with the fused TV + Adam step, whose atlas
    alternates between two buffers, for blocks of an odd number of steps (7) and a
    2-step tail: checkpoints and the final atlas."""
    torch = pytest.importorskip("torch")
    from photara_drender_amd import _native
    from photara_drender_amd.refine import refine_atlas

    sc, mats, masks, images, pairs = small_atlas(views=4)
    dev = torch.device("cuda:0")
    ctx = _native.Context(0)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    args = (t(np.full((64, 64, 3), 0.5, np.float32)), t(sc.positions), t(sc.uv), t(sc.indices.view(np.int32)),
            [t(im) for im in images], mats)
    kw = dict(visibility_masks=[t(m) for m in masks], seam_uv_pairs=t(pairs), steps=30, batch_size=2,
              learning_rate=1e-2, seam_polish_steps=4, tv_weight=1e-2, chart_ids=sc.chart_texels,
              joint_seam_weight=1e-2)
    a = refine_atlas(ctx, *args, use_graph=True, checkpoint_every=7, **kw)
    b = refine_atlas(ctx, *args, use_graph=False, checkpoint_every=7, **kw)
    assert a.checkpoints.shape == (4, 64, 64, 3)
    assert_bits(a.checkpoints.cpu().numpy(), b.checkpoints.cpu().numpy(), "checkpoints")
    assert_bits(a.color.cpu().numpy(), b.color.cpu().numpy(), "colour")
    assert_bits(a.loss_history, b.loss_history, "loss history")
    assert not torch.equal(a.checkpoints[0], a.checkpoints[1])
