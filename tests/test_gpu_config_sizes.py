"""Parity at the benchmark workloads' real sizes (GPU against the oracle, same
seeded inputs):

* 16 uneven views (512^2 .. 1920x1080, 15,163,392 px) of the 200,000-triangle
  heightfield with 8 attributes, batched forward + backward in one launch per
  kernel (src/rasterizer.cpp:373-829): ids, barycentrics, depth, derivatives,
  interpolated attributes, grad_raster and every view's position gradients
  bit-exact; the attribute gradient (float atomics) within 1e-4 * max + 1e-4
  relative, and bit-exact with PDR_FLAG_DETERMINISTIC (the oracle's ordered sum
  over the views' pixels in order);
* the 512^2 textured chain of the level-5 icosphere with the 1024^2 x 4
  texture: every forward output, grad_uv, grad_raster and position gradients
  bit-exact; texel and uv-attribute gradients within the same tolerance, and
  bit-exact with PDR_FLAG_DETERMINISTIC;
* the atlas refinement on the 1,000,000-triangle, 256-chart grid with its
  2048^2 atlas, 8 of its 1024x768 views, 10 steps of batch 4 and a 2-step seam
  polish over its 90,000 seam pairs (src/texture_refiner.cpp:174-454): colour
  and loss histories bit-exact.
"""
import ctypes

import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu


def bits_equal(a, b, what):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    diff = a.view(np.uint32) != b.view(np.uint32)
    assert not diff.any(), f"{what}: {int(diff.sum())} of {a.size} differ, max |d| {np.abs(a - b).max():.3g}"


def close_sum(gpu, ref, what, rel=1e-4):
    gpu, ref = np.asarray(gpu, np.float64), np.asarray(ref, np.float64)
    err = np.abs(gpu - ref)
    bound = rel * np.abs(ref).max() + rel * np.abs(ref)
    assert (err <= bound + 1e-30).all(), f"{what}: worst {err.max():.3g} vs scale {np.abs(ref).max():.3g}"


@pytest.fixture(scope="module")
def host_threads():
    import os

    oracle.set_num_threads(min(os.cpu_count() or 1, 16))
    return oracle.num_threads()


def _batched_16_views(ctx_flags):
    import torch

    from photara_drender_amd import _native

    scene = scenes.multiview_scene()
    dev = torch.device("cuda:0")
    V, T, A = scene.clips[0].shape[0], scene.indices.shape[0], scene.attributes.shape[1]
    pixels = [h * w for h, w in scene.resolutions]
    offs = np.concatenate([[0], np.cumsum(pixels)])
    P = int(offs[-1])
    idx = torch.tensor(scene.indices.view(np.int32), device=dev)
    attr = torch.tensor(scene.attributes, device=dev)
    clips = torch.tensor(np.stack(scene.clips), device=dev)
    target = torch.tensor(np.concatenate([t.reshape(-1, A) for t in scene.targets]), device=dev)
    raster = torch.empty((P, 4), device=dev)
    deriv = torch.empty((P, 4), device=dev)
    interp = torch.empty((P, A), device=dev)
    grad = torch.empty((P, A), device=dev)
    grad_raster = torch.empty((P, 4), device=dev)
    grad_attr = torch.empty((V, A), device=dev)
    grad_pos = torch.empty((len(pixels), V, 4), device=dev)
    n = len(pixels)
    rviews = (_native.RasterView * n)()
    iviews = (_native.InterpView * n)()
    for k, (h, w) in enumerate(scene.resolutions):
        o = int(offs[k])
        rv = rviews[k]
        rv.positions = clips[k].data_ptr()
        rv.width, rv.height = w, h
        rv.viewport = (ctypes.c_float * 4)(0, 0, 0, 0)
        rv.raster, rv.derivatives = raster[o].data_ptr(), deriv[o].data_ptr()
        rv.grad_raster, rv.grad_positions = grad_raster[o].data_ptr(), grad_pos[k].data_ptr()
        iv = iviews[k]
        iv.width, iv.height = w, h
        iv.raster, iv.out, iv.grad, iv.grad_raster = (raster[o].data_ptr(), interp[o].data_ptr(), grad[o].data_ptr(),
                                                      grad_raster[o].data_ptr())
    ctx = _native.Context(0, flags=ctx_flags)
    lib = _native.lib()
    st = torch.cuda.current_stream().cuda_stream
    _native.check(lib.pdr_rasterize_forward_batched(ctx.handle, rviews, n, V, idx.data_ptr(), T, 0, 1, st))
    _native.check(lib.pdr_interpolate_forward_batched(ctx.handle, iviews, n, attr.data_ptr(), V, A, idx.data_ptr(),
                                                      T, st))
    torch.sub(interp, target, out=grad)
    _native.check(lib.pdr_interpolate_backward_batched(ctx.handle, iviews, n, attr.data_ptr(), V, A, idx.data_ptr(),
                                                       T, grad_attr.data_ptr(), st))
    _native.check(lib.pdr_rasterize_backward_batched(ctx.handle, rviews, n, V, idx.data_ptr(), T, st))
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in dict(raster=raster, deriv=deriv, interp=interp, grad=grad,
                                                grad_raster=grad_raster, grad_attr=grad_attr,
                                                grad_pos=grad_pos).items()}
    return scene, offs, out


def test_sixteen_view_batched_chain_at_full_size(host_threads):
    from photara_drender_amd import _native

    scene, offs, out = _batched_16_views(0)
    A = scene.attributes.shape[1]
    rasters, grads = [], []
    for k, (h, w) in enumerate(scene.resolutions):
        sl = slice(int(offs[k]), int(offs[k + 1]))
        print(f"view {k} {w}x{h}", flush=True)
        ref_r, ref_d, _ = oracle.rasterize_forward(scene.clips[k], scene.indices, (h, w))
        bits_equal(out["raster"][sl], ref_r.reshape(-1, 4), f"view {k} raster")
        bits_equal(out["deriv"][sl], ref_d.reshape(-1, 4), f"view {k} derivatives")
        ref_i = oracle.interpolate_forward(scene.attributes, scene.indices, ref_r)
        bits_equal(out["interp"][sl], ref_i.reshape(-1, A), f"view {k} interpolated")
        g = (ref_i - scene.targets[k]).astype(np.float32)
        bits_equal(out["grad"][sl], g.reshape(-1, A), f"view {k} loss gradient")
        _, ref_gr = oracle.interpolate_backward(scene.attributes, scene.indices, ref_r, g)
        bits_equal(out["grad_raster"][sl], ref_gr.reshape(-1, 4), f"view {k} grad_raster")
        ref_gp = oracle.rasterize_backward(scene.clips[k], scene.indices, ref_r, ref_gr)
        bits_equal(out["grad_pos"][k], ref_gp, f"view {k} position gradient")
        rasters.append(ref_r.reshape(-1, 4))
        grads.append(g.reshape(-1, A))
    # the batched attribute gradient is one sum over every view's pixels in view order
    ref_ga, _ = oracle.interpolate_backward(scene.attributes, scene.indices, np.concatenate(rasters)[None],
                                            np.concatenate(grads)[None])
    close_sum(out["grad_attr"], ref_ga, "attribute gradient")
    _, _, det = _batched_16_views(_native.PDR_FLAG_DETERMINISTIC)
    bits_equal(det["grad_attr"], ref_ga, "attribute gradient (deterministic)")
    for key in ("raster", "deriv", "grad_raster", "grad_pos"):
        bits_equal(det[key], out[key], key + " (deterministic)")


def _chain(flags, s):
    import torch

    from photara_drender_amd import _native

    dev = torch.device("cuda:0")
    H, W = s.resolution
    th, tw, C = s.texture.shape
    V, T = s.clip.shape[0], s.indices.shape[0]
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    clip, idx, uv, tex, target = t(s.clip), t(s.indices.view(np.int32)), t(s.uv), t(s.texture), t(s.target)
    e = lambda *shape: torch.empty(shape, device=dev)
    raster, deriv, uvs, color = e(H, W, 4), e(H, W, 4), e(H, W, 2), e(H, W, C)
    g_color, g_tex, g_uvs, g_uv, g_raster, g_pos = e(H, W, C), e(th, tw, C), e(H, W, 2), e(V, 2), e(H, W, 4), e(V, 4)
    ctx = _native.Context(0, flags=flags)
    lib, ck = _native.lib(), _native.check
    st = torch.cuda.current_stream().cuda_stream
    ck(lib.pdr_rasterize_forward(ctx.handle, clip.data_ptr(), V, idx.data_ptr(), T, W, H, 0, 1, None,
                                 raster.data_ptr(), deriv.data_ptr(), None, st))
    ck(lib.pdr_interpolate_forward(ctx.handle, uv.data_ptr(), V, 2, idx.data_ptr(), T, raster.data_ptr(), W, H,
                                   uvs.data_ptr(), st))
    ck(lib.pdr_texture_forward(ctx.handle, tex.data_ptr(), tw, th, C, 0, uvs.data_ptr(), raster.data_ptr(), W, H,
                               color.data_ptr(), st))
    torch.sub(color, target, out=g_color).mul_(2.0)
    ck(lib.pdr_texture_backward(ctx.handle, tex.data_ptr(), tw, th, C, 0, uvs.data_ptr(), raster.data_ptr(), W, H,
                                g_color.data_ptr(), 1, g_tex.data_ptr(), g_uvs.data_ptr(), st))
    ck(lib.pdr_interpolate_backward(ctx.handle, uv.data_ptr(), V, 2, idx.data_ptr(), T, raster.data_ptr(), W, H,
                                    g_uvs.data_ptr(), g_uv.data_ptr(), g_raster.data_ptr(), st))
    ck(lib.pdr_rasterize_backward(ctx.handle, clip.data_ptr(), V, idx.data_ptr(), T, raster.data_ptr(), W, H,
                                  g_raster.data_ptr(), None, g_pos.data_ptr(), st))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in dict(raster=raster, deriv=deriv, uvs=uvs, color=color, g_color=g_color,
                                                g_tex=g_tex, g_uvs=g_uvs, g_uv=g_uv, g_raster=g_raster,
                                                g_pos=g_pos).items()}


def test_textured_chain_512_with_1024_texture(host_threads):
    from photara_drender_amd import _native

    s = scenes.chain_scene(level=5, res=(512, 512), tex_size=1024, seed=1)
    assert s.texture.shape == (1024, 1024, 4) and s.indices.shape[0] == 20480
    out = _chain(0, s)
    r, d, _ = oracle.rasterize_forward(s.clip, s.indices, s.resolution)
    bits_equal(out["raster"], r, "raster")
    bits_equal(out["deriv"], d, "derivatives")
    u = oracle.interpolate_forward(s.uv, s.indices, r)
    bits_equal(out["uvs"], u, "uv")
    c = oracle.texture_forward(s.texture, u, r)
    bits_equal(out["color"], c, "colour")
    gc = (2.0 * (c - s.target)).astype(np.float32)
    bits_equal(out["g_color"], gc, "loss gradient")
    gt, gu = oracle.texture_backward(s.texture, u, r, gc)
    bits_equal(out["g_uvs"], gu, "grad_uv")
    close_sum(out["g_tex"], gt, "texel gradient")
    ga, gr = oracle.interpolate_backward(s.uv, s.indices, r, gu)
    bits_equal(out["g_raster"], gr, "grad_raster")
    close_sum(out["g_uv"], ga, "uv attribute gradient")
    gp = oracle.rasterize_backward(s.clip, s.indices, r, gr)
    bits_equal(out["g_pos"], gp, "position gradient")
    det = _chain(_native.PDR_FLAG_DETERMINISTIC, s)
    bits_equal(det["g_tex"], gt, "texel gradient (deterministic)")
    bits_equal(det["g_uv"], ga, "uv attribute gradient (deterministic)")
    bits_equal(det["g_pos"], gp, "position gradient (deterministic)")


def test_atlas_refinement_at_full_mesh_and_atlas_size(host_threads):
    import photara_drender_amd as pdr

    sc = scenes.atlas_scene()  # 1000 x 500 quads, 16 x 16 charts, 2048^2 atlas
    assert sc.indices.shape[0] == 1_000_000 and sc.atlas_size == 2048
    size, views = (768, 1024), 8
    mats = scenes.atlas_views(sc, views, size, seed=3)
    masks = scenes.ellipse_masks(views, size, seed=3)
    rng = np.random.default_rng(3)
    images = []
    for k in range(views):
        clip = oracle.transform_positions(sc.positions, mats[k])
        r, _, _ = oracle.rasterize_forward(clip, sc.indices, size, want_deriv=False)
        surf = oracle.interpolate_forward(sc.surface, sc.indices, r)
        gt = scenes.ground_truth(surf.reshape(-1, 2)).reshape(size + (3,)) * (r[..., 3:4] != 0)
        images.append(np.clip(gt + rng.normal(0, 0.02, gt.shape), 0, 1).astype(np.float32))
    pairs = scenes.seam_uv_pairs(sc.indices, sc.uv, sc.vertex_remap, sc.face_chart_ids)
    assert pairs.shape[0] == 90000
    init = np.full((2048, 2048, 3), 0.5, np.float32)
    kw = dict(visibility_masks=masks, seam_uv_pairs=pairs, steps=10, batch_size=4, learning_rate=1e-2,
              seam_polish_steps=2)
    color, hist, seam, _, _ = pdr.TextureBaker().refine_texture(init, sc.positions, sc.uv, sc.indices, images, mats,
                                                                **kw)
    print("gpu refine done", flush=True)
    ref_c, ref_h, ref_s = oracle.refine(init, sc.positions, sc.uv, sc.indices, images, mats, **kw)
    bits_equal(hist, ref_h, "loss history")
    bits_equal(seam, ref_s, "seam loss history")
    bits_equal(color, ref_c, "atlas")
    assert hist[-1] < hist[0]
