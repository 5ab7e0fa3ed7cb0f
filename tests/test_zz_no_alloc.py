"""pdr_rasterize_forward_noalloc: the rasterize forward without device
allocations (scratch in a fixed per-GPU array) against pdr_rasterize_forward on
the same device arrays: raster and derivatives bit-exact on two calls in a row,
the second smaller than the first; free device memory unchanged across a call;
calls past the array's 32 MiB and host arrays are rejected before any GPU work.
"""
import ctypes

import numpy as np
import pytest

from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu


def _scene(res, level, seed, big):
    """Icosphere view plus `big` screen-sized triangles (the queued large-triangle path)."""
    s = scenes.chain_scene(level=level, res=res, tex_size=8, seed=seed)
    clip, idx = s.clip, s.indices
    if big:
        rng = np.random.default_rng(seed)
        v = rng.uniform(-1.2, 1.2, (3 * big, 4)).astype(np.float32)
        v[:, 2] = rng.uniform(-0.5, 0.9, 3 * big)
        v[:, 3] = 1.0
        extra = clip.shape[0] + np.arange(3 * big, dtype=np.uint32).reshape(big, 3)
        clip, idx = np.concatenate([clip, v]), np.concatenate([idx, extra])
    return np.ascontiguousarray(clip), np.ascontiguousarray(idx), res


def _run(entry, ctx, clip, idx, res, vp=None, deriv=True):
    import torch

    from photara_drender_amd import _native

    dev = torch.device("cuda:0")
    h, w = res
    c, i = torch.tensor(clip, device=dev), torch.tensor(idx.view(np.int32), device=dev)
    r = torch.full((h, w, 4), -7.0, device=dev)
    d = torch.full((h, w, 4), -7.0, device=dev)
    vpa = None if vp is None else (ctypes.c_float * 4)(*vp)
    out = (ctypes.c_float * 4)()
    st = torch.cuda.current_stream().cuda_stream
    _native.check(getattr(_native.lib(), entry)(ctx.handle, c.data_ptr(), clip.shape[0], i.data_ptr(), idx.shape[0],
                                                w, h, 0, 1 if deriv else 0, vpa, r.data_ptr(),
                                                d.data_ptr() if deriv else None, out, st))
    torch.cuda.synchronize()
    return r.cpu().numpy(), d.cpu().numpy(), list(out)


def _bits(a, b, what):
    assert a.shape == b.shape
    diff = a.view(np.uint32) != b.view(np.uint32)
    assert not diff.any(), f"{what}: {int(diff.sum())} values differ"


@pytest.fixture(scope="module")
def ctx():
    from photara_drender_amd import _native

    return _native.Context(0)


def test_noalloc_forward_equals_the_allocating_forward_twice(ctx):
    cases = [(_scene((600, 800), 5, 3, 3), None), (_scene((150, 200), 3, 4, 1), (10.0, 5.0, 160.0, 120.0))]
    for (clip, idx, res), vp in cases:  # the second call is smaller than the first
        got = _run("pdr_rasterize_forward_noalloc", ctx, clip, idx, res, vp)
        want = _run("pdr_rasterize_forward", ctx, clip, idx, res, vp)
        _bits(got[0], want[0], f"raster {res}")
        _bits(got[1], want[1], f"derivatives {res}")
        assert got[2] == want[2]
        assert (got[0][..., 3] > 0).sum() > 1000
    # without derivatives the derivative array is not touched
    clip, idx, res = cases[1][0]
    r, d, _ = _run("pdr_rasterize_forward_noalloc", ctx, clip, idx, res, None, deriv=False)
    _bits(r, _run("pdr_rasterize_forward", ctx, clip, idx, res, None, deriv=False)[0], "raster without derivatives")
    assert (d == -7.0).all()


def test_noalloc_forward_allocates_nothing(ctx):
    import torch

    from photara_drender_amd import _native

    clip, idx, res = _scene((512, 512), 4, 5, 2)
    dev = torch.device("cuda:0")
    h, w = res
    c, i = torch.tensor(clip, device=dev), torch.tensor(idx.view(np.int32), device=dev)
    r, d = torch.empty((h, w, 4), device=dev), torch.empty((h, w, 4), device=dev)
    st = torch.cuda.current_stream().cuda_stream
    lib = _native.lib()

    def call():
        _native.check(lib.pdr_rasterize_forward_noalloc(ctx.handle, c.data_ptr(), clip.shape[0], i.data_ptr(),
                                                        idx.shape[0], w, h, 0, 1, None, r.data_ptr(), d.data_ptr(),
                                                        None, st))
        torch.cuda.synchronize()

    call()  # code objects loaded
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        call()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0


def test_noalloc_forward_rejects_oversized_and_host_calls(ctx):
    import torch

    from photara_drender_amd import _native

    dev = torch.device("cuda:0")
    c = torch.zeros((3, 4), device=dev)
    i = torch.zeros((1, 3), dtype=torch.int32, device=dev)
    r = torch.zeros((1, 1, 4), device=dev)  # never written: the calls fail before any GPU work
    lib = _native.lib()
    # 4096 x 4096 needs 128 MiB of keys: over the 32 MiB array
    rc = lib.pdr_rasterize_forward_noalloc(ctx.handle, c.data_ptr(), 3, i.data_ptr(), 1, 4096, 4096, 0, 0, None,
                                           r.data_ptr(), None, None, None)
    assert rc == 3 and b"limit is 33554432" in lib.pdr_last_error()
    hc = np.zeros((3, 4), np.float32)
    hi = np.zeros((1, 3), np.uint32)
    hr = np.zeros((1, 1, 4), np.float32)
    rc = lib.pdr_rasterize_forward_noalloc(ctx.handle, hc.ctypes.data, 3, hi.ctypes.data, 1, 1, 1, 0, 0, None,
                                           hr.ctypes.data, None, None, None)
    assert rc == 1 and b"device memory only" in lib.pdr_last_error()
