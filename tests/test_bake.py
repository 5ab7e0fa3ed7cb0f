"""Texture bake (TextureBaker::bake, src/texture_baker.cpp:156-361;
shaders/project_texture.hlsl:117-247).

CPU: the oracle's restatement pinned by closed forms -
* an orthographic view of a flat quad whose atlas maps texel (i, j) onto pixel
  centre (i + 0.5, j + 0.5): every blend mode returns the image itself;
* occlusion through the shadow map, masks, minimum view cosine;
* best_view keeps the highest-confidence view, weighted_average the
  confidence-weighted mean, softmax the exp-weighted mean with the
  reference's running maximum;
* suggest_softmax_scale = 20 * (closest vertex-camera distance)^2.
GPU: pdr_bake against the oracle on a multi-chart mesh and several views -
coverage, face ids, source views, validity bit-exact; colours and confidences
bit-exact for best_view and weighted_average (same float expressions, -ffp-contract=off,
correctly rounded division and sqrt), within 1e-5 relative for softmax (device
expf vs host exp).
"""
import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes

ORTHO = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, -0.1, 0], [0, 0, 0, 1]], np.float32)


def quad(z=0.0, size=1.0):
    pos = np.array([[-size, -size, z], [size, -size, z], [size, size, z], [-size, size, z]], np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return pos, nrm, uv, idx


@pytest.mark.parametrize("mode", ["best_view", "weighted_average", "softmax"])
def test_orthographic_quad_returns_the_image(mode):
    pos, nrm, uv, idx = quad()
    rng = np.random.default_rng(0)
    img = rng.uniform(0, 1, (24, 32, 3)).astype(np.float32)
    color, conf, src, valid, cov, face = oracle.bake(pos, nrm, uv, idx, [img], ORTHO[None], [[0, 0, 1000.0]],
                                                     resolution=(24, 32), blend_mode=mode)
    assert cov.all() and valid.all() and (src == 0).all()
    np.testing.assert_allclose(color[..., :3], img, atol=2e-5)
    assert (color[..., 3] == 1).all()
    assert set(np.unique(face)) <= {0, 1}
    if mode != "softmax":
        np.testing.assert_allclose(conf, 1.0, atol=1e-6)  # visibility * alpha * cos * weight


def test_occluded_surface_gets_no_samples_and_masks_veto():
    front, nrm, uv, idx = quad(z=0.5, size=0.5)
    back, _, _, _ = quad(z=0.0, size=1.0)
    img = np.full((32, 32, 3), 0.25, np.float32)
    # bake the back quad while the front quad casts a shadow: merge the meshes, atlas only the back quad's uv
    pos = np.concatenate([back, front])
    nrms = np.concatenate([nrm, nrm])
    uvs = np.concatenate([uv, np.full((4, 2), -5.0, np.float32)])  # front quad outside the atlas
    idxs = np.concatenate([idx, idx + 4]).astype(np.uint32)
    color, conf, src, valid, cov, _ = oracle.bake(pos, nrms, uvs, idxs, [img], ORTHO[None], [[0, 0, 1000.0]],
                                                  resolution=(32, 32), blend_mode="best_view", pcf_radius=0)
    assert cov.all()
    # texels under the front quad (x, y in [-0.5, 0.5]) are occluded, the others see the image
    c = (np.arange(32) + 0.5) / 32 * 2 - 1
    inside = (np.abs(c)[None, :] < 0.45) & (np.abs(c)[:, None] < 0.45)
    outside = (np.abs(c)[None, :] > 0.55) | (np.abs(c)[:, None] > 0.55)
    assert not valid[inside].any() and valid[outside].all()
    mask = np.zeros((32, 32), np.float32)
    _, _, _, valid2, _, _ = oracle.bake(pos, nrms, uvs, idxs, [img], ORTHO[None], [[0, 0, 1000.0]], [mask],
                                        resolution=(32, 32))
    assert not valid2.any()
    _, _, _, valid3, _, _ = oracle.bake(pos, nrms, uvs, idxs, [img], ORTHO[None], [[0, 0, 1000.0]], [mask],
                                        resolution=(32, 32), mask_floor=0.5)
    assert valid3[outside].all()


def test_blend_modes_combine_views_in_closed_form():
    pos, nrm, uv, idx = quad()
    a = np.full((16, 16, 3), 0.2, np.float32)
    b = np.full((16, 16, 3), 0.8, np.float32)
    mats = np.stack([ORTHO, ORTHO])
    cams = [[0, 0, 1000.0], [0, 0, 1000.0]]
    kw = dict(resolution=(16, 16), weights=[1.0, 3.0])
    col, conf, src, *_ = oracle.bake(pos, nrm, uv, idx, [a, b], mats, cams, blend_mode="best_view", **kw)
    np.testing.assert_allclose(col[..., :3], 0.8, atol=1e-6)
    assert (src == 1).all()
    col, conf, src, *_ = oracle.bake(pos, nrm, uv, idx, [a, b], mats, cams, blend_mode="weighted_average", **kw)
    cos = conf / 4.0
    np.testing.assert_allclose(col[..., :3], (0.2 * 1 + 0.8 * 3) / 4, atol=1e-6)
    assert (cos > 0.99).all()
    scale = 2.0
    col, conf, src, *_ = oracle.bake(pos, nrm, uv, idx, [a, b], mats, cams, blend_mode="softmax", softmax_scale=scale,
                                     **kw)
    # exponent_k = weight_k * scale * cos / d^2 (clamped to [0, 64])
    d2 = np.float32(1000.0) ** 2
    e = np.array([1.0, 3.0])[:, None, None] * scale * cos[None] / d2
    w = np.exp(e - e.max(0))
    np.testing.assert_allclose(col[..., 0], (0.2 * w[0] + 0.8 * w[1]) / w.sum(0), rtol=1e-5)
    assert (src == 1).all()


def test_suggest_softmax_scale_closed_form():
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(50, 3)).astype(np.float32)
    cams = rng.normal(size=(4, 3)).astype(np.float32) * 5
    d2 = ((pos[:, None, :].astype(np.float64) - cams[None].astype(np.float64)) ** 2).sum(-1).min()
    assert oracle.suggest_softmax_scale(pos, cams) == pytest.approx(20 * d2, rel=1e-6)


def bake_scene(views=4, seed=5, size=(48, 64)):
    sc = scenes.atlas_scene(nx=40, ny=20, charts_x=4, charts_y=2, atlas_size=64, gutter=2, seed=seed)
    mats = scenes.atlas_views(sc, views, size, seed=seed)
    rng = np.random.default_rng(seed)
    images = [rng.uniform(0, 1, size + (3 + (k % 2),)).astype(np.float32) for k in range(views)]
    masks = [rng.uniform(0, 1, size).astype(np.float32) if k % 2 else None for k in range(views)]
    cams = np.stack([scenes.camera_center(m) for m in mats])
    nrm = scenes.vertex_normals(sc.positions, sc.indices)
    return sc, nrm, mats, cams, images, masks


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["best_view", "weighted_average", "softmax"])
def test_gpu_bake_matches_oracle(mode):
    import photara_drender_amd as pdr

    sc, nrm, mats, cams, images, masks = bake_scene()
    baker = pdr.TextureBaker()
    kw = dict(resolution=(64, 64), blend_mode=mode, mask_floor=0.1)
    g = baker.bake(sc.positions, nrm, sc.uv, sc.indices, images, mats, cams, masks, None,
                   visibility_mode="shadow_map", **kw)
    r = oracle.bake(sc.positions, nrm, sc.uv, sc.indices, images, mats, cams, masks, **kw)
    color, conf, src, valid, cov, used = g
    assert used is False
    np.testing.assert_array_equal(cov, r[4])
    assert valid.sum() > 100
    if mode == "softmax":
        np.testing.assert_allclose(color, r[0], rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(conf, r[1], rtol=1e-5, atol=1e-6)
        agree = (src == r[2]).mean()
        assert agree > 0.999  # ties of equal exponents can flip with expf rounding
    else:
        np.testing.assert_array_equal(color.view(np.uint32), r[0].view(np.uint32))
        np.testing.assert_array_equal(conf.view(np.uint32), r[1].view(np.uint32))
        np.testing.assert_array_equal(src, r[2])
        np.testing.assert_array_equal(valid, r[3])


@pytest.mark.gpu
@pytest.mark.parametrize("pcf", [0, 2, 3])
def test_gpu_bake_pcf_radii_match_oracle(pcf):
    """Radii 0 and 2 run kernels with the radius fixed at compile time, 3 the run-time loop."""
    import photara_drender_amd as pdr

    sc, nrm, mats, cams, images, masks = bake_scene()
    kw = dict(resolution=(64, 64), blend_mode="weighted_average", mask_floor=0.1, pcf_radius=pcf)
    g = pdr.TextureBaker().bake(sc.positions, nrm, sc.uv, sc.indices, images, mats, cams, masks, None,
                                visibility_mode="shadow_map", **kw)
    r = oracle.bake(sc.positions, nrm, sc.uv, sc.indices, images, mats, cams, masks, **kw)
    assert g[3].sum() > 100
    np.testing.assert_array_equal(g[0].view(np.uint32), r[0].view(np.uint32))
    np.testing.assert_array_equal(g[1].view(np.uint32), r[1].view(np.uint32))
    np.testing.assert_array_equal(g[2], r[2])


@pytest.mark.gpu
@pytest.mark.parametrize("pcf", [1, 2])
def test_gpu_bake_at_unaligned_sizes_matches_oracle(pcf):
    """Views of 45 x 61 and a 61 x 67 atlas: sides that are not multiples of 4, so the
    raster keys' 4 x 4-pixel tiles (the shadow maps the projection reads) are padded."""
    import photara_drender_amd as pdr

    sc, nrm, mats, cams, images, masks = bake_scene(size=(45, 61))
    kw = dict(resolution=(61, 67), blend_mode="weighted_average", mask_floor=0.1, pcf_radius=pcf)
    g = pdr.TextureBaker().bake(sc.positions, nrm, sc.uv, sc.indices, images, mats, cams, masks, None,
                                visibility_mode="shadow_map", **kw)
    r = oracle.bake(sc.positions, nrm, sc.uv, sc.indices, images, mats, cams, masks, **kw)
    assert g[3].sum() > 100
    np.testing.assert_array_equal(g[4], r[4])
    np.testing.assert_array_equal(g[0].view(np.uint32), r[0].view(np.uint32))
    np.testing.assert_array_equal(g[1].view(np.uint32), r[1].view(np.uint32))
    np.testing.assert_array_equal(g[2], r[2])


@pytest.mark.gpu
def test_gpu_bake_with_autotuned_raster_matches_oracle():
    """The shadow maps are the raster's candidate keys; under PDR_FLAG_AUTOTUNE their
    tuning runs clear the keys between candidates, and the bake stays bit-exact."""
    from photara_drender_amd import _native
    from photara_drender_amd.bake import bake_atlas

    sc, nrm, mats, cams, images, masks = bake_scene()
    ctx = _native.Context(0, flags=_native.PDR_FLAG_AUTOTUNE | _native.PDR_FLAG_VALIDATE_DEVICE_INPUTS)
    kw = dict(resolution=(64, 64), blend_mode="weighted_average", mask_floor=0.1)
    for _ in range(2):  # the first call tunes, the second uses the cached choice
        g = bake_atlas(ctx, sc.positions, nrm, sc.uv, sc.indices, images, mats, cams, masks, **kw)
        r = oracle.bake(sc.positions, nrm, sc.uv, sc.indices, images, mats, cams, masks, **kw)
        np.testing.assert_array_equal(g.color.cpu().numpy().view(np.uint32), r[0].view(np.uint32))
        np.testing.assert_array_equal(g.confidence.cpu().numpy().view(np.uint32), r[1].view(np.uint32))


@pytest.mark.gpu
def test_gpu_bake_validation_messages():
    import photara_drender_amd as pdr

    pos, nrm, uv, idx = quad()
    img = np.zeros((8, 8, 3), np.float32)
    baker = pdr.TextureBaker()
    with pytest.raises(ValueError, match="normals must have shape"):
        baker.bake(pos, nrm[:3], uv, idx, [img], ORTHO[None], [[0, 0, 1.0]])
    with pytest.raises(RuntimeError, match="ray-query visibility was requested"):
        baker.bake(pos, nrm, uv, idx, [img], ORTHO[None], [[0, 0, 1.0]], allow_visibility_fallback=False)
    with pytest.raises(ValueError, match="pcf_radius exceeds 4"):
        baker.bake(pos, nrm, uv, idx, [img], ORTHO[None], [[0, 0, 1.0]], pcf_radius=5)
    with pytest.raises(ValueError, match="blend_mode must be"):
        baker.bake(pos, nrm, uv, idx, [img], ORTHO[None], [[0, 0, 1.0]], blend_mode="max")
    with pytest.raises(IndexError, match="invalid vertex index"):
        baker.bake(pos, nrm, uv, np.array([[0, 1, 7]], np.uint32), [img], ORTHO[None], [[0, 0, 1.0]])
