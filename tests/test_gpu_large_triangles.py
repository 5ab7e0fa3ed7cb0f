"""Screen-filling triangles against the oracle: the forward's queued 16-row slabs
(more than 16,384 candidate pixels per triangle) and the backward's large boxes
(more than 65,536 pixels), as the reference's own test triangle
(tests/native_tests.cpp:125-157) scaled to 512^2 and 1024^2, perspective
triangles, and small triangles in front of them.

Bars:
* forward (ids, barycentrics, depth, derivatives): bit-exact;
* backward, default: the large boxes are summed by a fixed workgroup tree, not
  in the reference's row-major order (rasterize_backward.hlsl:88-151), so their
  vertices' gradients are within 1e-4 * max|oracle| + 1e-4 * |oracle|; vertices
  of small triangles only are bit-exact;
* backward with PDR_FLAG_DETERMINISTIC: every box in the reference's order,
  bit-exact.
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TRIANGLE = np.array([[-0.75, -0.75, 0.0, 1.0], [0.75, -0.75, 0.0, 1.0], [0.0, 0.75, 0.0, 1.0]], np.float32)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def bits_equal(a, b, what):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape
    diff = a.view(np.uint32) != b.view(np.uint32)
    assert not diff.any(), f"{what}: {int(diff.sum())} values differ, max |d| {np.abs(a - b).max():.3g}"


def close_sum(gpu, ref, rel=1e-4):
    gpu, ref = np.asarray(gpu, np.float64), np.asarray(ref, np.float64)
    err = np.abs(gpu - ref)
    bound = rel * np.abs(ref).max() + rel * np.abs(ref)
    assert (err <= bound + 1e-30).all(), f"worst {err.max():.3g} vs scale {np.abs(ref).max():.3g}"


def small_occluders(rng, n, first):
    """n small triangles in front (depth -0.9 .. -0.5) scattered over the screen."""
    c = rng.uniform(-0.9, 0.9, (n, 1, 2))
    v = np.zeros((n, 3, 4), np.float32)
    v[..., :2] = c + rng.uniform(-0.04, 0.04, (n, 3, 2))
    v[..., 2] = rng.uniform(-0.9, -0.5, (n, 3))
    v[..., 3] = 1.0
    idx = first + np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return v.reshape(-1, 4), idx


def case(name):
    """(clip positions, indices, (H, W), number of leading vertices that belong to large triangles)"""
    rng = np.random.default_rng(sum(map(ord, name)))
    if name.startswith("native"):
        n = int(name.split("_")[1])
        return TRIANGLE, np.array([[0, 1, 2]], np.uint32), (n, n), 3
    if name == "perspective_quad_1024":
        # a receding quad (w from 0.7 to 2.4) covering most of the screen, x,y scaled by w
        ndc = np.array([[-1.1, -1.05], [1.05, -1.1], [1.1, 1.05], [-1.05, 1.1]])
        w = np.array([0.7, 0.9, 2.4, 1.8])
        z = np.array([-0.2, 0.1, 0.6, 0.3]) * w
        big = np.concatenate([ndc * w[:, None], z[:, None], w[:, None]], 1).astype(np.float32)
        bidx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
        sv, sidx = small_occluders(rng, 150, 4)
        return np.concatenate([big, sv]), np.concatenate([bidx, sidx]), (1024, 1024), 4
    if name == "overlapping_fan_512":
        # five large perspective triangles crossing each other at different depths
        k = 5
        ang = np.linspace(0, 2 * np.pi, k, endpoint=False)
        big = []
        for i, a in enumerate(ang):
            w = rng.uniform(0.8, 1.6, 3)
            pts = np.array([[np.cos(a), np.sin(a)], [np.cos(a + 2.3), np.sin(a + 2.3)],
                            [np.cos(a + 4.1), np.sin(a + 4.1)]]) * 1.05
            z = rng.uniform(-0.4, 0.8, 3)
            big.append(np.concatenate([pts * w[:, None], (z * w)[:, None], w[:, None]], 1))
        big = np.concatenate(big).astype(np.float32)
        bidx = np.arange(3 * k, dtype=np.uint32).reshape(k, 3)
        sv, sidx = small_occluders(rng, 80, 3 * k)
        return np.concatenate([big, sv]), np.concatenate([bidx, sidx]), (512, 512), 3 * k
    raise KeyError(name)


CASES = ["native_512", "native_1024", "perspective_quad_1024", "overlapping_fan_512"]


@pytest.fixture(scope="module")
def det_ctx():
    from photara_drender_amd import _native

    return _native.Context(0, flags=_native.PDR_FLAG_DETERMINISTIC | _native.PDR_FLAG_VALIDATE_DEVICE_INPUTS)


@pytest.mark.parametrize("name", CASES)
def test_large_triangles_forward_bit_exact(rasterizer, name):
    pos, idx, res, _ = case(name)
    raster, deriv = rasterizer.forward(pos, idx, res)
    ref_r, ref_d, _ = oracle.rasterize_forward(pos, idx, res)
    bits_equal(raster, ref_r, "raster")
    bits_equal(deriv, ref_d, "derivatives")
    # the large triangles name far more pixels than one wave walks (16,384 candidates)
    ids, counts = np.unique(ref_r[..., 3], return_counts=True)
    assert counts[ids > 0].max() > 65536


@pytest.mark.parametrize("name", CASES)
def test_large_triangles_backward(rasterizer, det_ctx, name):
    from photara_drender_amd import _native

    pos, idx, res, n_big = case(name)
    ref_r, _, _ = oracle.rasterize_forward(pos, idx, res, 0, False)
    rng = np.random.default_rng(7)
    grad = rng.standard_normal(ref_r.shape).astype(np.float32)
    grad[rng.random(res) < 0.1, :2] = 0.0
    ref = oracle.rasterize_backward(pos, idx, ref_r, grad)
    # default: tree order inside the large boxes
    g = rasterizer.backward(pos, idx, ref_r, grad)
    assert (g[:, 2] == 0).all()
    close_sum(g, ref)
    bits_equal(g[n_big:], ref[n_big:], "small-triangle vertices")
    # PDR_FLAG_DETERMINISTIC: the reference's row-major order everywhere
    h, w = res
    out = np.empty_like(ref)
    pos_c, idx_c = np.ascontiguousarray(pos), np.ascontiguousarray(idx)
    _native.check(_native.lib().pdr_rasterize_backward(det_ctx.handle, _p(pos_c), pos_c.shape[0], _p(idx_c),
                                                       idx_c.shape[0], _p(ref_r), w, h, _p(grad), None, _p(out),
                                                       None))
    bits_equal(out, ref, "ordered backward")


@pytest.mark.parametrize("name", ["perspective_quad_1024", "overlapping_fan_512"])
def test_large_work_queue_under_autotuning_and_repeated_calls(name):
    """The large-work queue's counters are zero between calls without a fill
    launch: the resolve pass (forward) and the vertex reduce (backward) clear them
    after the queue's consumer.  The autotuner's first call runs every candidate
    launch shape back to back on the same queue, and later calls reuse it: each
    forward / backward pair must still match the oracle, call after call."""
    from photara_drender_amd import _native

    pos, idx, res, n_big = case(name)
    h, w = res
    ref_r, ref_d, _ = oracle.rasterize_forward(pos, idx, res)
    rng = np.random.default_rng(11)
    grad = rng.standard_normal(ref_r.shape).astype(np.float32)
    ref_g = oracle.rasterize_backward(pos, idx, ref_r, grad)
    ctx = _native.Context(0, flags=_native.PDR_FLAG_AUTOTUNE | _native.PDR_FLAG_DETERMINISTIC)
    lib = _native.lib()
    pos_c, idx_c = np.ascontiguousarray(pos), np.ascontiguousarray(idx)
    for call in range(3):
        r = np.empty((h, w, 4), np.float32)
        d = np.empty((h, w, 4), np.float32)
        _native.check(lib.pdr_rasterize_forward(ctx.handle, _p(pos_c), pos_c.shape[0], _p(idx_c), idx_c.shape[0], w,
                                                h, 0, 1, None, _p(r), _p(d), None, None))
        bits_equal(r, ref_r, f"raster, call {call}")
        bits_equal(d, ref_d, f"derivatives, call {call}")
        g = np.empty_like(ref_g)
        _native.check(lib.pdr_rasterize_backward(ctx.handle, _p(pos_c), pos_c.shape[0], _p(idx_c), idx_c.shape[0],
                                                 _p(r), w, h, _p(grad), None, _p(g), None))
        bits_equal(g, ref_g, f"position gradients, call {call}")
    kinds = sorted(line.split()[0] for line in ctx.autotune_export().splitlines())
    assert kinds == ["backward", "forward_out"]
