"""Autograd check of the full chain (rasterize -> interpolate -> texture) through the
torch.autograd.Functions of photara_drender_amd.torch (python/photara_drender/torch.py:
17-241), on the autograd-check workload: a 2,048-triangle UV sphere (32 segments x
33 rings), a 128x128 view and a 256x256 RGBA texture, loss = sum(color * weights).

Gradients w.r.t. clip positions, UV attributes and texels are compared
* with the oracle's FP32 chain backward (the reference's algorithm restated):
  positions bit-exact (per-pixel grad_raster is exact and the position sums run in
  the reference's order); UV attributes and texels, whose sums are scattered with
  float atomics, within |gpu - oracle| <= 1e-4 * max|oracle| + 1e-4 * |oracle|;
* with central finite differences (h = 1e-4) of the oracle's FP64 chain with frozen
  visibility (triangle ids of the FP32 raster): texels and UV attributes within
  1e-3 relative, positions within 2e-3 relative (the reference's backward uses
  1 / (area +- 1e-6)).  Bilinear weights are piecewise linear in uv: a sample whose
  one-sided difference quotients disagree by more than 1e-3 (the step crosses a
  texel boundary for some pixel) is skipped.
"""
import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu

RES = (128, 128)


def chain_inputs(seed=7):
    pos, uv, idx = scenes.uv_sphere(32, 33)
    assert idx.shape[0] == 2048
    m = scenes.world_to_clip(np.array([0.4, 0.9, 2.6]), 1.0)
    clip = scenes.to_clip(pos, m).astype(np.float32)
    rng = np.random.default_rng(seed)
    tex = rng.uniform(0, 1, (256, 256, 4)).astype(np.float32)
    wgt = rng.standard_normal(RES + (4,)).astype(np.float32)
    return clip, uv.astype(np.float32), idx, tex, wgt


def oracle_chain_grads(clip, uv, idx, tex, wgt):
    r, _, _ = oracle.rasterize_forward(clip, idx, RES)
    uvs = oracle.interpolate_forward(uv, idx, r)
    g_tex, g_uvs = oracle.texture_backward(tex, uvs, r, wgt, "clamp", True)
    g_uv, g_raster = oracle.interpolate_backward(uv, idx, r, g_uvs)
    g_pos = oracle.rasterize_backward(clip, idx, r, g_raster)
    return r, g_pos, g_uv, g_tex


def loss_f64(clip, uv, tex, idx, ids, wgt):
    r = oracle.raster_fixed_f64(clip, idx, ids)
    uvs = oracle.interpolate_forward_f64(uv, idx, r)
    color = oracle.texture_forward_f64(tex, uvs, r)
    return float((color * wgt).sum())


def close_sum(gpu, ref, rel=1e-4):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = np.abs(ref).max()
    bad = np.abs(gpu - ref) > rel * scale + rel * np.abs(ref) + 1e-30
    assert not bad.any(), f"{bad.sum()} of {bad.size} exceed tolerance"


def test_chain_gradients_match_oracle_and_finite_differences():
    import torch

    import photara_drender_amd as pdr
    from photara_drender_amd import torch as pdt

    clip, uv, idx, tex, wgt = chain_inputs()
    dev = torch.device("cuda:0")
    rz = pdr.Rasterizer()
    pos_t = torch.tensor(clip, device=dev, requires_grad=True)
    uv_t = torch.tensor(uv, device=dev, requires_grad=True)
    tex_t = torch.tensor(tex, device=dev, requires_grad=True)
    idx_t = torch.tensor(idx.astype(np.int64), device=dev)
    raster, _ = pdt.rasterize(pos_t, idx_t, RES, rasterizer=rz)
    uvs = pdt.interpolate(uv_t, raster, idx_t, rasterizer=rz)
    color = pdt.texture(tex_t, uvs, raster, rasterizer=rz)
    (color * torch.tensor(wgt, device=dev)).sum().backward()
    g_pos, g_uv, g_tex = (t.grad.cpu().numpy() for t in (pos_t, uv_t, tex_t))

    r, o_pos, o_uv, o_tex = oracle_chain_grads(clip, uv, idx, tex, wgt)
    np.testing.assert_array_equal(raster.detach().cpu().numpy().view(np.uint32), r.view(np.uint32))
    np.testing.assert_array_equal(g_pos.view(np.uint32), o_pos.view(np.uint32))
    close_sum(g_uv, o_uv)
    close_sum(g_tex, o_tex)

    # central differences of the FP64 chain, visibility frozen at the FP32 triangle ids
    h = 1e-4
    rng = np.random.default_rng(3)
    c64, uv64, tex64 = clip.astype(np.float64), uv.astype(np.float64), tex.astype(np.float64)
    f = lambda c, u, t: loss_f64(c, u, t, idx, r, wgt)

    touched = np.argwhere(np.abs(o_tex).sum(-1) > 0)
    for ty, tx in touched[rng.choice(len(touched), 8, replace=False)]:
        ch = int(rng.integers(0, 4))
        tp, tm = tex64.copy(), tex64.copy()
        tp[ty, tx, ch] += h
        tm[ty, tx, ch] -= h
        fd = (f(c64, uv64, tp) - f(c64, uv64, tm)) / (2 * h)
        assert abs(g_tex[ty, tx, ch] - fd) <= 1e-3 * max(abs(fd), 1e-2), (ty, tx, ch, g_tex[ty, tx, ch], fd)

    visible = np.unique(idx[np.unique(r[..., 3][r[..., 3] > 0]).astype(np.int64) - 1])
    checked = 0
    for v in rng.choice(visible, 24, replace=False):
        for c in (0, 1):
            up, um = uv64.copy(), uv64.copy()
            up[v, c] += h
            um[v, c] -= h
            f0, fp, fm = f(c64, uv64, tex64), f(c64, up, tex64), f(c64, um, tex64)
            right, left = (fp - f0) / h, (f0 - fm) / h
            if abs(right - left) > 1e-3 * max(abs(right), abs(left), 1.0):
                continue  # the step crosses a texel boundary for some pixel (bilinear kink)
            fd = (fp - fm) / (2 * h)
            assert abs(g_uv[v, c] - fd) <= 1e-3 * max(abs(fd), 1.0), (v, c, g_uv[v, c], fd)
            checked += 1
    assert checked >= 12

    checked = 0
    for v in rng.choice(visible, 24, replace=False):
        for c in (0, 1, 3):
            cp, cm = c64.copy(), c64.copy()
            cp[v, c] += h
            cm[v, c] -= h
            f0, fp, fm = f(c64, uv64, tex64), f(cp, uv64, tex64), f(cm, uv64, tex64)
            right, left = (fp - f0) / h, (f0 - fm) / h
            if abs(right - left) > 1e-3 * max(abs(right), abs(left), 1.0):
                continue
            fd = (fp - fm) / (2 * h)
            assert abs(g_pos[v, c] - fd) <= 2e-3 * max(abs(fd), 1.0), (v, c, g_pos[v, c], fd)
            checked += 1
    assert checked >= 8
    assert (g_pos[:, 2] == 0).all()
