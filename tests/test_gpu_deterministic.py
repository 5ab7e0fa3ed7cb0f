"""PDR_FLAG_DETERMINISTIC: the attribute and texel gradients as ordered gathers.

With the flag set, pdr_interpolate_backward(_batched) and pdr_texture_backward
sum each vertex's / texel's contributions in ascending pixel order (tap order
within a pixel), the order of the reference's per-vertex pixel loop and of its
texel adjacency lists.  The oracle's gradients are those same ordered sums, so
the device results must equal them bit for bit, where the default atomic path
only matches them to a tolerance."""
import ctypes

import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _bits_equal(a, b):
    np.testing.assert_array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.fixture(scope="module")
def det_ctx():
    from photara_drender_amd import _native

    return _native.Context(0, flags=_native.PDR_FLAG_DETERMINISTIC | 1)


@pytest.mark.parametrize("attrs", [1, 3, 8, 11])
def test_attribute_gradient_is_the_ordered_sum(det_ctx, attrs):
    from photara_drender_amd import _native

    s = scenes.chain_scene(level=4, res=(160, 192), tex_size=32, seed=attrs)
    raster, _, _ = oracle.rasterize_forward(s.clip, s.indices, s.resolution)
    rng = np.random.default_rng(attrs)
    V = s.clip.shape[0]
    attr = rng.standard_normal((V, attrs)).astype(np.float32)
    h, w = raster.shape[:2]
    grad = rng.standard_normal((h, w, attrs)).astype(np.float32)
    idx = np.ascontiguousarray(s.indices, dtype=np.uint32)
    ga = np.empty_like(attr)
    gr = np.empty((h, w, 4), np.float32)
    _native.check(_native.lib().pdr_interpolate_backward(det_ctx.handle, _p(attr), V, attrs, _p(idx), idx.shape[0],
                                                         _p(raster), w, h, _p(grad), _p(ga), _p(gr), None))
    ref_a, ref_r = oracle.interpolate_backward(attr, idx, raster, grad)
    _bits_equal(gr, ref_r)
    _bits_equal(ga, ref_a)


def test_batched_attribute_gradient_is_the_ordered_sum_over_views(det_ctx):
    """Two views of one size: the views' pixels in order are the rows of the
    two rasters stacked, on which the oracle's ordered sum runs."""
    import torch

    from photara_drender_amd import _native

    s = scenes.chain_scene(level=4, res=(96, 128), tex_size=32, seed=3)
    rng = np.random.default_rng(7)
    rasters = []
    for k in range(2):
        clip = np.array(s.clip, dtype=np.float32)
        clip[:, 0] += np.float32(0.15 * k) * clip[:, 3]  # the second view shifted by 0.15 in NDC x
        rasters.append(oracle.rasterize_forward(clip, s.indices, s.resolution)[0])
    V, A = s.clip.shape[0], 8
    attr = rng.standard_normal((V, A)).astype(np.float32)
    h, w = rasters[0].shape[:2]
    grads = [rng.standard_normal((h, w, A)).astype(np.float32) for _ in range(2)]
    dev = torch.device("cuda:0")
    t_attr = torch.tensor(attr, device=dev)
    t_idx = torch.tensor(s.indices.astype(np.int32), device=dev)
    t_r = [torch.tensor(r, device=dev) for r in rasters]
    t_g = [torch.tensor(g, device=dev) for g in grads]
    t_gr = [torch.empty((h, w, 4), device=dev) for _ in range(2)]
    t_ga = torch.empty((V, A), device=dev)
    views = (_native.InterpView * 2)()
    for k in range(2):
        views[k].raster = t_r[k].data_ptr()
        views[k].grad = t_g[k].data_ptr()
        views[k].grad_raster = t_gr[k].data_ptr()
        views[k].width, views[k].height = w, h
    _native.check(_native.lib().pdr_interpolate_backward_batched(det_ctx.handle, views, 2, t_attr.data_ptr(), V, A,
                                                                 t_idx.data_ptr(), s.indices.shape[0],
                                                                 t_ga.data_ptr(), None))
    torch.cuda.synchronize()
    ref_a, _ = oracle.interpolate_backward(attr, s.indices, np.concatenate(rasters, 0), np.concatenate(grads, 0))
    _bits_equal(t_ga.cpu().numpy(), ref_a)


@pytest.mark.parametrize("mode", ["clamp", "wrap"])
@pytest.mark.parametrize("channels", [1, 3, 4, 5])
def test_texel_gradient_is_the_ordered_sum(det_ctx, mode, channels):
    from photara_drender_amd import _native

    s = scenes.chain_scene(level=4, res=(150, 170), tex_size=48, channels=min(channels, 4), seed=channels)
    raster, _, _ = oracle.rasterize_forward(s.clip, s.indices, s.resolution)
    uv = oracle.interpolate_forward(s.uv, s.indices, raster)
    if mode != "clamp":
        uv = (uv * 3.0 - 1.25).astype(np.float32)
    rng = np.random.default_rng(channels)
    tex = rng.uniform(0, 1, (48, 48, channels)).astype(np.float32)
    h, w = raster.shape[:2]
    grad = rng.standard_normal((h, w, channels)).astype(np.float32)
    gt = np.empty_like(tex)
    gu = np.empty((h, w, 2), np.float32)
    m = {"clamp": 0, "wrap": 1}[mode]
    _native.check(_native.lib().pdr_texture_backward(det_ctx.handle, _p(tex), 48, 48, channels, m, _p(uv), _p(raster),
                                                     w, h, _p(grad), 1, _p(gt), _p(gu), None))
    ref_t, ref_u = oracle.texture_backward(tex, uv, raster, grad, mode, True)
    _bits_equal(gu, ref_u)
    _bits_equal(gt, ref_t)


def test_empty_raster_gives_zero_gradients(det_ctx):
    """No covered pixel: every entry is a sentinel, every vertex / texel sums nothing."""
    from photara_drender_amd import _native

    s = scenes.chain_scene(level=3, res=(40, 56), tex_size=16, seed=1)
    V, A, h, w = s.clip.shape[0], 3, 40, 56
    raster = np.zeros((h, w, 4), np.float32)
    rng = np.random.default_rng(0)
    attr = rng.standard_normal((V, A)).astype(np.float32)
    grad = rng.standard_normal((h, w, A)).astype(np.float32)
    idx = np.ascontiguousarray(s.indices, dtype=np.uint32)
    ga = np.full_like(attr, np.nan)
    gr = np.empty((h, w, 4), np.float32)
    _native.check(_native.lib().pdr_interpolate_backward(det_ctx.handle, _p(attr), V, A, _p(idx), idx.shape[0],
                                                         _p(raster), w, h, _p(grad), _p(ga), _p(gr), None))
    assert (ga == 0).all()
    tex = rng.uniform(0, 1, (16, 16, 3)).astype(np.float32)
    uv = rng.uniform(0, 1, (h, w, 2)).astype(np.float32)
    gt = np.full_like(tex, np.nan)
    gu = np.empty((h, w, 2), np.float32)
    g3 = rng.standard_normal((h, w, 3)).astype(np.float32)
    _native.check(_native.lib().pdr_texture_backward(det_ctx.handle, _p(tex), 16, 16, 3, 0, _p(uv), _p(raster), w, h,
                                                     _p(g3), 1, _p(gt), _p(gu), None))
    assert (gt == 0).all()
