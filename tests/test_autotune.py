"""Autotuner of the raster's launch parameters (PDR_FLAG_AUTOTUNE; project-specified).

* The forward with outputs chooses between the tiled path (bins of 32x32
  raster tiles, LDS keys) at bin depths of 2, 4 and 8 tiles per triangle and
  the triangle-parallel path (global keys) with its walk limit, wave count and
  traversal order (index order, or each view's nearer end first); the backward
  chooses its walk limit and wave count.
* Raster forward and backward results are bit-identical under the defaults,
  under the autotuned choice, and under every candidate imported into the cache
  (the forward's winner is order-free; both backward paths sum each triangle in
  the reference's order).
* The cache exports one line per workload class, imports back (a cached class is
  not re-tuned), and rejects malformed lines.
"""
import ctypes

import numpy as np
import pytest

from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu

FORWARD = [(1, d, 16384.0, 4096, 0) for d in (2, 4, 8)] + \
    [(0, 4, w, n, o) for o in (0, 1) for w in (4096.0, 16384.0, 65536.0) for n in (2048, 4096, 8192)]
BACKWARD = [(w, n) for w in (256.0, 512.0, 1024.0, 2048.0) for n in (2048, 4096, 8192, 16384)] * 2


def run_views(ctx, sc):
    import torch

    from photara_drender_amd import _native

    dev = torch.device("cuda:0")
    V, T = sc.clips[0].shape[0], sc.indices.shape[0]
    idx = torch.tensor(sc.indices.view(np.int32), device=dev)
    n = len(sc.clips)
    clips = [torch.tensor(c, device=dev) for c in sc.clips]
    rasters = [torch.empty((h, w, 4), device=dev) for h, w in sc.resolutions]
    derivs = [torch.empty((h, w, 4), device=dev) for h, w in sc.resolutions]
    rng = np.random.default_rng(3)
    grs = [torch.tensor(rng.normal(size=(h, w, 4)).astype(np.float32), device=dev) for h, w in sc.resolutions]
    gpos = [torch.empty((V, 4), device=dev) for _ in range(n)]
    rv = (_native.RasterView * n)()
    for k, (h, w) in enumerate(sc.resolutions):
        rv[k].positions, rv[k].width, rv[k].height = clips[k].data_ptr(), w, h
        rv[k].viewport = (ctypes.c_float * 4)(0, 0, 0, 0)
        rv[k].raster, rv[k].derivatives = rasters[k].data_ptr(), derivs[k].data_ptr()
        rv[k].grad_raster, rv[k].grad_positions = grs[k].data_ptr(), gpos[k].data_ptr()
    lib = _native.lib()
    s = torch.cuda.current_stream().cuda_stream
    _native.check(lib.pdr_rasterize_forward_batched(ctx.handle, rv, n, V, idx.data_ptr(), T, 0, 1, s))
    _native.check(lib.pdr_rasterize_backward_batched(ctx.handle, rv, n, V, idx.data_ptr(), T, s))
    torch.cuda.synchronize()
    return [np.concatenate([a.cpu().numpy().ravel().view(np.uint32) for a in group])
            for group in (rasters, derivs, gpos)]


def scene():
    return scenes.multiview_scene(nx=120, ny=80, views=3, attributes=2, seed=11,
                                  sizes=[(96, 128), (200, 150), (64, 64)])


def test_autotuned_and_every_candidate_give_identical_results():
    from photara_drender_amd import _native

    sc = scene()
    ref = run_views(_native.Context(0, flags=0), sc)
    tuned_ctx = _native.Context(0, flags=_native.PDR_FLAG_AUTOTUNE)
    got = run_views(tuned_ctx, sc)
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a, b)
    lines = tuned_ctx.autotune_export().splitlines()
    kinds = sorted(line.split()[0] for line in lines)
    assert kinds == ["backward", "forward_out"], lines
    keys = {line.split()[0]: " ".join(line.split()[:4]) for line in lines}
    for (tiled, depth, ww, fw, order), (flat, bw) in zip(FORWARD, BACKWARD):
        ctx = _native.Context(0, flags=_native.PDR_FLAG_AUTOTUNE)
        ctx.autotune_import(f"{keys['forward_out']} {ww} {fw} 512 4096 {tiled} {depth} {order}\n"
                            f"{keys['backward']} 16384 4096 {flat} {bw} 1 4\n")
        got = run_views(ctx, sc)
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b)
        assert ctx.autotune_export() == f"{keys['backward']} 16384 4096 {flat:g} {bw} 1 4 0\n" \
                                        f"{keys['forward_out']} {ww:g} {fw} 512 4096 {tiled} {depth} {order}\n"  # not re-tuned


def test_autotune_cache_rejects_malformed_lines():
    from photara_drender_amd import _native

    ctx = _native.Context(0, flags=0)
    with pytest.raises(ValueError, match="malformed"):
        ctx.autotune_import("forward 1 2\n")
    with pytest.raises(ValueError, match="malformed"):
        ctx.autotune_import("sideways 1 2 3 4096 4096 512 4096 1 4\n")
    with pytest.raises(ValueError, match="malformed"):
        ctx.autotune_import("forward_out 1 2 3 4096 4096 512 4096 2 4\n")  # tiled is 0 or 1
    with pytest.raises(ValueError, match="malformed"):
        ctx.autotune_import("forward_out 1 2 3 4096 4096 512 4096 0 4 2\n")  # order is 0 or 1
    with pytest.raises(ValueError, match="malformed"):
        ctx.autotune_import("forward_out 1 2 3 4096 4096 512 4096 0 4 1 7\n")  # one field too many
    ctx.autotune_import("forward 1 12 20 4096 2048 512 4096 0 4\n\n")  # without the order: index order
    assert ctx.autotune_export() == "forward 1 12 20 4096 2048 512 4096 0 4 0\n"
    ctx.autotune_import("forward 1 12 20 4096 2048 512 4096 0 4 1\n")
    assert ctx.autotune_export() == "forward 1 12 20 4096 2048 512 4096 0 4 1\n"
    ctx.autotune_clear()
    assert ctx.autotune_export() == ""
