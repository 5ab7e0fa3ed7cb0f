"""The atlas losses and loops (python/photara_drender/atlas.py:65-257 semantics).

CPU: the masked photometric loss and the chart-aware total variation against
closed forms and brute-force loops over texel pairs; the validation messages.
GPU:
* the seam loss against a float64 restatement of clamp-mode bilinear sampling
  (texture_forward.hlsl: t = uv * size - 0.5, corner order 00, 10, 01, 11);
* optimize_texture_atlas: its first Adam step moves every texel by
  lr * g / (|g| + 1e-8) (torch.optim.Adam's first bias-corrected step) with g the
  float64 autograd gradient of photometric + tv + seam (+ prior) restated on
  the CPU; the raster cache gives bit-identical atlases and histories;
* the refiner's fused extended step (photometric + chart TV + joint seam,
  pdr_refine with tv_weight / joint_seam_weight) the same way, for the first
  Adam step of atlas_adam_update.hlsl with its bias corrections.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from photara_drender_amd import scenes  # noqa: E402
from photara_drender_amd.atlas import atlas_total_variation, masked_charbonnier_loss  # noqa: E402


def test_masked_charbonnier_closed_form():
    rng = np.random.default_rng(1)
    r = rng.uniform(0, 1, (9, 11, 3))
    t = rng.uniform(0, 1, (9, 11, 3))
    m = (rng.uniform(size=(9, 11, 1)) < 0.6).astype(np.float64) * rng.uniform(0.2, 1, (9, 11, 1))
    eps = 1e-3
    want = (np.sqrt((r - t) ** 2 + eps ** 2) * m).sum() / (max(m.sum(), 1.0) * 3)
    got = masked_charbonnier_loss(torch.tensor(r), torch.tensor(t), torch.tensor(m), epsilon=eps)
    assert float(got) == pytest.approx(want, rel=1e-12)
    # no valid pixel: the denominator is clamped to 1 * channels
    z = masked_charbonnier_loss(torch.tensor(r), torch.tensor(t), torch.zeros(9, 11, 1), epsilon=eps)
    assert float(z) == 0.0
    with pytest.raises(ValueError, match="rendered and target must have the same shape"):
        masked_charbonnier_loss(torch.zeros(2, 2, 3), torch.zeros(2, 2, 4), torch.zeros(2, 2, 1))
    with pytest.raises(ValueError, match=r"valid_mask must have shape \[height, width, 1\]"):
        masked_charbonnier_loss(torch.zeros(2, 2, 3), torch.zeros(2, 2, 3), torch.zeros(2, 2))


def test_total_variation_brute_force():
    rng = np.random.default_rng(2)
    h, w, c = 7, 9, 3
    tex = rng.uniform(0, 1, (h, w, c))
    valid = rng.uniform(size=(h, w)) < 0.8
    chart = (np.arange(w)[None, :] // 4 + 3 * (np.arange(h)[:, None] // 3)).astype(np.int64)
    eps = 1e-4
    total, pairs = 0.0, 0
    for y in range(h):
        for x in range(w):
            for dy, dx in ((0, 1), (1, 0)):
                yy, xx = y + dy, x + dx
                if yy < h and xx < w and valid[y, x] and valid[yy, xx] and chart[y, x] == chart[yy, xx]:
                    total += np.sqrt((tex[yy, xx] - tex[y, x]) ** 2 + eps ** 2).sum()
                    pairs += 1
    got = atlas_total_variation(torch.tensor(tex), valid_mask=torch.tensor(valid), chart_ids=torch.tensor(chart),
                                epsilon=eps)
    assert float(got) == pytest.approx(total / (pairs * c), rel=1e-12)
    # without charts and mask every neighbour pair counts
    n_all = (h * (w - 1) + (h - 1) * w)
    full = (np.sqrt(np.diff(tex, axis=1) ** 2 + eps ** 2).sum() + np.sqrt(np.diff(tex, axis=0) ** 2 + eps ** 2).sum())
    assert float(atlas_total_variation(torch.tensor(tex), epsilon=eps)) == pytest.approx(full / (n_all * c), rel=1e-12)
    with pytest.raises(ValueError, match=r"texture_values must have shape \[height, width, channels\]"):
        atlas_total_variation(torch.zeros(3, 3))
    with pytest.raises(ValueError, match="valid_mask must have shape"):
        atlas_total_variation(torch.zeros(3, 3, 3), valid_mask=torch.ones(3, 2))
    with pytest.raises(ValueError, match="chart_ids must have shape"):
        atlas_total_variation(torch.zeros(3, 3, 3), chart_ids=torch.ones(2, 3))


# ---------------------------------------------------------------- GPU


def bilinear64(atlas, uv):
    """Clamp-mode bilinear samples [N, C] of a float64 atlas tensor at float32 uv
    [N, 2]: positions and weights rounded as the kernels round them, the sum in
    float64 (differentiable in the atlas)."""
    h, w = atlas.shape[:2]
    uv = np.asarray(uv, np.float32).reshape(-1, 2)
    tx = uv[:, 0] * np.float32(w) - np.float32(0.5)
    ty = uv[:, 1] * np.float32(h) - np.float32(0.5)
    flx, fly = np.floor(tx), np.floor(ty)
    fx, fy = (tx - flx).astype(np.float32), (ty - fly).astype(np.float32)
    one = np.float32(1)
    x0, y0 = np.clip(flx.astype(np.int64), 0, w - 1), np.clip(fly.astype(np.int64), 0, h - 1)
    x1, y1 = np.clip(flx.astype(np.int64) + 1, 0, w - 1), np.clip(fly.astype(np.int64) + 1, 0, h - 1)
    wts = [(one - fx) * (one - fy), fx * (one - fy), (one - fx) * fy, fx * fy]
    taps = [(y0, x0), (y0, x1), (y1, x0), (y1, x1)]
    out = 0
    for wk, (ty_, tx_) in zip(wts, taps):
        out = out + torch.tensor(wk.astype(np.float64))[:, None] * atlas[torch.tensor(ty_), torch.tensor(tx_)]
    return out


def small_scene(views=2, size=(72, 96), seed=8):
    sc = scenes.atlas_scene(nx=40, ny=20, charts_x=4, charts_y=2, atlas_size=64, gutter=2, seed=seed)
    mats = scenes.atlas_views(sc, views, size, seed=seed)
    masks = scenes.ellipse_masks(views, size, seed=seed)
    rng = np.random.default_rng(seed)
    images, uvs, rasters, clips = [], [], [], []
    for k in range(views):
        clip = oracle.transform_positions(sc.positions, mats[k])
        r, _, _ = oracle.rasterize_forward(clip, sc.indices, size, want_deriv=False)
        images.append(rng.uniform(0, 1, size + (3,)).astype(np.float32))
        uvs.append(oracle.interpolate_forward(sc.uv, sc.indices, r))
        rasters.append(r)
        clips.append(clip)
    pairs = scenes.seam_uv_pairs(sc.indices, sc.uv, sc.vertex_remap, sc.face_chart_ids)
    return sc, mats, masks, images, uvs, rasters, clips, pairs


def first_adam_step_check(p0, p1, g64, lr, what):
    """p1 = p0 - lr * g / (|g| + 1e-8) for the first bias-corrected Adam step."""
    g = g64.detach().numpy()
    strong = np.abs(g) > 1e-5
    assert strong.sum() > 100
    want = p0.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    err = np.abs(p1.astype(np.float64) - want)[strong]
    assert err.max() < 1e-6, f"{what}: worst {err.max():.3g}"
    untouched = g == 0
    assert np.array_equal(p1[untouched], p0[untouched]), f"{what}: texels without gradient moved"


@pytest.mark.gpu
def test_seam_loss_matches_bilinear_restatement(rasterizer):
    from photara_drender_amd.atlas import seam_consistency_loss

    sc, *_, pairs = small_scene()
    rng = np.random.default_rng(3)
    atlas = rng.uniform(0, 1, (64, 64, 3)).astype(np.float32)
    dev = torch.device("cuda:0")
    got = seam_consistency_loss(torch.tensor(atlas, device=dev), torch.tensor(pairs, device=dev),
                                rasterizer=rasterizer)
    a64 = torch.tensor(atlas.astype(np.float64))
    c = bilinear64(a64, pairs.reshape(-1, 2)).reshape(-1, 2, 3)
    want = torch.sqrt((c[:, 0] - c[:, 1]) ** 2 + 1e-8).mean()
    assert float(got) == pytest.approx(float(want), rel=1e-5)
    with pytest.raises(ValueError, match=r"seam_uv_pairs must have shape \[seam_count, 2, 2\]"):
        seam_consistency_loss(torch.zeros(64, 64, 3, device=dev), torch.zeros(4, 4, device=dev),
                              rasterizer=rasterizer)
    empty = seam_consistency_loss(torch.zeros(64, 64, 3, device=dev, requires_grad=True),
                                  torch.zeros(0, 2, 2, device=dev), rasterizer=rasterizer)
    assert float(empty.detach()) == 0.0 and empty.requires_grad


@pytest.mark.gpu
def test_optimize_texture_atlas_first_step_and_cache(rasterizer):
    from photara_drender_amd.atlas import AtlasOptimizationOptions, optimize_texture_atlas, seam_consistency_loss

    sc, mats, masks, images, uvs, rasters, clips, pairs = small_scene()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(4)
    init = rng.uniform(0.2, 0.8, (64, 64, 3)).astype(np.float32)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    args = (t(init), t(np.stack(clips)), t(sc.indices.astype(np.int64)), t(sc.uv), [t(im) for im in images])
    chart = sc.chart_texels
    kw = dict(rasterizer=rasterizer, visibility_masks=[t(m) for m in masks], atlas_valid_mask=t(chart >= 0),
              chart_ids=t(chart), seam_uv_pairs=t(pairs))
    opt = AtlasOptimizationOptions(steps=1, learning_rate=1e-2, total_variation_weight=1e-2, seam_weight=1e-1,
                                   initial_prior_weight=0.5)
    p1, hist = optimize_texture_atlas(*args, options=opt, **kw)
    assert len(hist) == 1
    # float64 autograd of the same objective (view 0; the prior has zero gradient at the start)
    a64 = torch.tensor(init.astype(np.float64), requires_grad=True)
    covered = rasters[0][..., 3].reshape(-1) > 0
    m = (covered * masks[0].reshape(-1)).astype(np.float64)
    rendered = bilinear64(a64, uvs[0].reshape(-1, 2))
    d = rendered - torch.tensor(images[0].reshape(-1, 3).astype(np.float64))
    photo = (torch.sqrt(d ** 2 + 1e-6) * torch.tensor(m)[:, None]).sum() / (max(m.sum(), 1.0) * 3)
    tv = atlas_total_variation(a64, valid_mask=torch.tensor(chart >= 0), chart_ids=torch.tensor(chart))
    c = bilinear64(a64, pairs.reshape(-1, 2)).reshape(-1, 2, 3)
    seam = torch.sqrt((c[:, 0] - c[:, 1]) ** 2 + 1e-8).mean()
    loss64 = photo + 1e-2 * tv + 1e-1 * seam
    loss64.backward()
    assert hist[0] == pytest.approx(float(loss64), rel=1e-5)
    first_adam_step_check(init, p1.cpu().numpy(), a64.grad, 1e-2, "optimize_texture_atlas")
    # the raster cache changes nothing (rasterising fixed positions is deterministic;
    # the texel gradients' float atomics are made ordered for the comparison)
    import photara_drender_amd as pdr
    from photara_drender_amd import _native

    det = pdr.Rasterizer()
    _native.check(_native.lib().pdr_context_set_flags(det.context_handle, _native.PDR_FLAG_DETERMINISTIC |
                                                      _native.PDR_FLAG_VALIDATE_DEVICE_INPUTS))
    kw["rasterizer"] = det
    opt4 = AtlasOptimizationOptions(steps=4, learning_rate=1e-2, total_variation_weight=1e-2, seam_weight=1e-1)
    a, ha = optimize_texture_atlas(*args, options=opt4, **kw)
    b, hb = optimize_texture_atlas(*args, options=AtlasOptimizationOptions(**{**opt4.__dict__,
                                                                              "cache_view_rasters": True}), **kw)
    assert torch.equal(a, b) and ha == hb
    # round robin: step 1 is view 1 (its loss differs from a view-0-only run's)
    same_view = optimize_texture_atlas(args[0], args[1][:1].repeat(2, 1, 1), args[2], args[3],
                                       [args[4][0], args[4][0]], options=opt4,
                                       **{**kw, "visibility_masks": [kw["visibility_masks"][0]] * 2})[1]
    assert same_view[0] == ha[0] and same_view[1] != ha[1]
    del seam_consistency_loss


@pytest.mark.gpu
def test_refine_extended_first_step_matches_autograd():
    """pdr_refine with chart TV and the joint seam term: one step, both views in the
    batch; the fused TV + Adam kernel uses the hardware square root and
    reciprocal, so the step is checked to 1e-6 (not bitwise)."""
    from photara_drender_amd import _native
    from photara_drender_amd.refine import refine_atlas

    sc, mats, masks, images, uvs, rasters, clips, pairs = small_scene()
    rng = np.random.default_rng(5)
    init = rng.uniform(0.2, 0.8, (64, 64, 3)).astype(np.float32)
    chart = sc.chart_texels
    tv_w, seam_w, lr = 3e-2, 2e-1, 1e-2
    res = refine_atlas(_native.Context(0), init, sc.positions, sc.uv, sc.indices, images, mats, masks, None, pairs,
                       steps=1, batch_size=2, learning_rate=lr, seam_polish_steps=0, tv_weight=tv_w,
                       chart_ids=chart, joint_seam_weight=seam_w, use_graph=False)
    a64 = torch.tensor(init.astype(np.float64), requires_grad=True)
    num, den = 0, 0.0
    for k in range(2):  # texture_refiner.cpp:316-330: one normalisation over the batch's mask sums
        m = ((rasters[k][..., 3].reshape(-1) != 0) * np.clip(masks[k].reshape(-1), 0, 1)).astype(np.float64)
        d = bilinear64(a64, uvs[k].reshape(-1, 2)) - torch.tensor(images[k].reshape(-1, 3).astype(np.float64))
        num = num + (torch.sqrt(d ** 2 + 1e-6) * torch.tensor(m)[:, None]).sum()
        den += m.sum()
    photo = num / (3 * den)
    tv = atlas_total_variation(a64, valid_mask=torch.tensor(chart >= 0), chart_ids=torch.tensor(chart))
    c = bilinear64(a64, pairs.reshape(-1, 2)).reshape(-1, 2, 3)
    seam = torch.sqrt((c[:, 0] - c[:, 1]) ** 2 + 1e-8).mean()
    (photo + tv_w * tv + seam_w * seam).backward()
    assert res.loss_history[0] == pytest.approx(float(photo), rel=1e-5)  # the history holds the photometric loss
    first_adam_step_check(init, np.asarray(res.color), a64.grad, lr, "refine extended step")
