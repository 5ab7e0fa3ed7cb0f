"""The tiled forward (triangles binned into 32x32 raster tiles, LDS depth keys,
tile-local resolve) forced through the autotuner's cache at bin depths of 2,
4 and 8 tiles per triangle: raster and derivatives bit-exact against the CPU
oracle, on the 512^2 icosphere, overlapping perspective soups (with culling),
screen-filling triangles over small occluders (the unbinned triangles' global
keys seed their tiles), and the 16-view, 200,000-triangle batched workload
against the default triangle-parallel path."""
import ctypes

import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu


def _lg(x):
    b = 0
    while b < 63 and (1 << b) < x:
        b += 1
    return b


def tiled_ctx(views, T, pixels, depth):
    """A context whose cache holds the tiled forward for this workload class."""
    from photara_drender_amd import _native

    ctx = _native.Context(0, flags=_native.PDR_FLAG_AUTOTUNE | _native.PDR_FLAG_VALIDATE_DEVICE_INPUTS)
    ctx.autotune_import(f"forward_out {views} {_lg(T)} {_lg(pixels)} 16384 4096 512 4096 1 {depth}\n")
    return ctx


def forward(ctx, pos, idx, res, cull=0):
    from photara_drender_amd import _native

    h, w = res
    pos, idx = np.ascontiguousarray(pos, np.float32), np.ascontiguousarray(idx, np.uint32)
    r = np.empty((h, w, 4), np.float32)
    d = np.empty((h, w, 4), np.float32)
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    _native.check(_native.lib().pdr_rasterize_forward(ctx.handle, p(pos), pos.shape[0], p(idx), idx.shape[0], w, h,
                                                      cull, 1, None, p(r), p(d), None, None))
    return r, d


def bits(a, b, what):
    diff = np.ascontiguousarray(a).view(np.uint32) != np.ascontiguousarray(b).view(np.uint32)
    assert not diff.any(), f"{what}: {int(diff.sum())} values differ"


def cases():
    s = scenes.chain_scene(level=5, res=(512, 512), tex_size=8, seed=1)
    yield "icosphere_512", s.clip, s.indices, (512, 512), 0
    rng = np.random.default_rng(5)
    for cull in (0, 1, 2):
        n = 600
        v = rng.uniform(-1.3, 1.3, (3 * n, 4)).astype(np.float32)
        v[:, 2] = rng.uniform(-1.2, 1.2, 3 * n)
        v[:, 3] = rng.uniform(0.5, 2.0, 3 * n)
        v[:, :2] *= v[:, 3:4]
        yield f"soup_cull{cull}", v, np.arange(3 * n, dtype=np.uint32).reshape(n, 3), (181, 233), cull
    big = np.array([[-1.1, -1.05, -0.2, 1.0], [1.05, -1.1, 0.1, 1.0], [0.2, 1.1, 0.6, 1.0]], np.float32)
    c = rng.uniform(-0.9, 0.9, (300, 1, 2))
    small = np.zeros((300, 3, 4), np.float32)
    small[..., :2] = c + rng.uniform(-0.03, 0.03, (300, 3, 2))
    small[..., 2] = rng.uniform(-0.9, -0.5, (300, 3))
    small[..., 3] = 1.0
    pos = np.concatenate([big, small.reshape(-1, 4)])
    yield "screen_filling_640x480", pos, np.arange(pos.shape[0], dtype=np.uint32).reshape(-1, 3), (480, 640), 0


@pytest.mark.parametrize("depth", [2, 4, 8])
def test_tiled_forward_bit_exact(depth):
    for name, pos, idx, res, cull in cases():
        ctx = tiled_ctx(1, idx.shape[0], res[0] * res[1], depth)
        r, d = forward(ctx, pos, idx, res, cull)
        ref_r, ref_d, _ = oracle.rasterize_forward(pos, idx, res, cull, True)
        bits(r, ref_r, f"{name} raster (depth {depth})")
        bits(d, ref_d, f"{name} derivatives (depth {depth})")
        assert "forward_out" in ctx.autotune_export() and ctx.autotune_export().split()[8] == "1"  # the tiled field


def test_tiled_forward_sixteen_views_equals_default_path():
    import torch

    from photara_drender_amd import _native

    sc = scenes.multiview_scene()
    dev = torch.device("cuda:0")
    V, T = sc.clips[0].shape[0], sc.indices.shape[0]
    idx = torch.tensor(sc.indices.view(np.int32), device=dev)
    clips = [torch.tensor(c, device=dev) for c in sc.clips]
    P = sum(h * w for h, w in sc.resolutions)
    out = []
    for ctx in (_native.Context(0, flags=0), tiled_ctx(len(clips), T, P, 4)):
        rasters = [torch.empty((h, w, 4), device=dev) for h, w in sc.resolutions]
        derivs = [torch.empty((h, w, 4), device=dev) for h, w in sc.resolutions]
        rv = (_native.RasterView * len(clips))()
        for k, (h, w) in enumerate(sc.resolutions):
            rv[k].positions, rv[k].width, rv[k].height = clips[k].data_ptr(), w, h
            rv[k].viewport = (ctypes.c_float * 4)(0, 0, 0, 0)
            rv[k].raster, rv[k].derivatives = rasters[k].data_ptr(), derivs[k].data_ptr()
        _native.check(_native.lib().pdr_rasterize_forward_batched(ctx.handle, rv, len(clips), V, idx.data_ptr(), T,
                                                                  0, 1, torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        out.append([np.concatenate([a.cpu().numpy().ravel() for a in g]) for g in (rasters, derivs)])
    bits(out[1][0], out[0][0], "16-view raster")
    bits(out[1][1], out[0][1], "16-view derivatives")
