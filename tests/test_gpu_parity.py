"""GPU parity: the HIP kernels (through the reference's numpy API and through the
zero-copy torch path) against the CPU oracle on the same seeded inputs.

Bars (written per assertion):
* forward outputs (triangle ids, barycentrics, depth, derivatives, interpolated
  attributes, texture samples) and the per-pixel backward outputs (grad_raster of
  interpolation, grad_uv of texturing) are bit-exact;
* scattered gradient sums (positions, attributes, texels) differ from the
  oracle only in summation order: |gpu - oracle| <= 1e-4 * max|oracle| + 1e-4 * |oracle|.
"""
import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes

pytestmark = pytest.mark.gpu

TRIANGLE = np.array([[-0.75, -0.75, 0.0, 1.0], [0.75, -0.75, 0.0, 1.0], [0.0, 0.75, 0.0, 1.0]], np.float32)
ONE = np.array([[0, 1, 2]], np.uint32)


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape
    diff = a.view(np.uint32) != b.view(np.uint32)
    if diff.any():
        where = np.argwhere(diff)[:5]
        raise AssertionError(f"{diff.sum()} values differ, first at {where.tolist()}: "
                             f"{a[tuple(where[0])]} vs {b[tuple(where[0])]}")


def close_sum(gpu, ref, rel=1e-4):
    """Atomic-order tolerance for scattered sums."""
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = np.abs(ref).max() if ref.size else 0.0
    err = np.abs(gpu - ref)
    bound = rel * scale + rel * np.abs(ref)
    bad = err > bound + 1e-30
    assert not bad.any(), f"{bad.sum()} of {bad.size} exceed tolerance; worst {err.max()} vs scale {scale}"


def soup(seed, n=300, perspective=True):
    """Random overlapping triangles, some with w <= 0 (dropped) and degenerate ones."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1.3, 1.3, size=(3 * n, 4)).astype(np.float32)
    v[:, 2] = rng.uniform(-1.2, 1.2, size=3 * n)
    v[:, 3] = rng.uniform(0.5, 2.0, size=3 * n) if perspective else 1.0
    v[:, :2] *= v[:, 3:4]
    v[rng.choice(3 * n, 5, replace=False), 3] = -0.5
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    idx[7] = [idx[7, 0], idx[7, 0], idx[7, 1]]  # degenerate
    # shared edges with equal depth: lower index must win
    v[30:33] = [[-0.5, -0.5, 0.25, 1.0], [0.5, -0.5, 0.25, 1.0], [0.0, 0.5, 0.25, 1.0]]
    v[33:36] = [[0.5, -0.5, 0.25, 1.0], [-0.5, -0.5, 0.25, 1.0], [0.0, 0.5, 0.25, 1.0]]
    return v, idx


FORWARD_CASES = {
    "native_triangle_32": (TRIANGLE, ONE, (32, 32), None, 0),
    "native_viewport_32": (TRIANGLE, ONE, (32, 32), (8.0, 4.0, 16.0, 20.0), 0),
    "smoke_16x20": (TRIANGLE, ONE, (16, 20), None, 0),
    "smoke_viewport_20x24": (TRIANGLE, ONE, (20, 24), (4.0, 3.0, 16.0, 12.0), 0),
    "nvdiffrast_64": (np.array([[-0.75, -0.75, 0.1, 1.0], [0.75, -0.75, 0.2, 1.0], [0.0, 0.75, 0.3, 1.0]],
                               np.float32), ONE, (64, 64), None, 0),
}


def tiny_soup(seed, n=400):
    """Sub-pixel, epsilon-scale and sliver triangles whose epsilon-inclusive coverage
    reaches past their bounding boxes (rasterize_forward.hlsl:103-113)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.05, 1.05, size=(n, 1, 2))
    size = np.where(rng.random((n, 1, 1)) < 0.5, 1e-5, 2e-3)
    v = np.zeros((n, 3, 4), np.float32)
    v[..., :2] = c + rng.uniform(-1, 1, size=(n, 3, 2)) * size
    v[::3, 2, :2] = 0.5 * (v[::3, 0, :2] + v[::3, 1, :2]) + 1e-7  # slivers
    v[..., 2] = rng.uniform(-1, 1, size=(n, 3))
    v[..., 3] = 1.0
    return v.reshape(-1, 4), np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def _case(name):
    if name in FORWARD_CASES:
        return FORWARD_CASES[name]
    if name.startswith("tiny_"):
        v, idx = tiny_soup(int(name.split("_")[1]))
        return v, idx, (61, 53), None, 0
    if name == "icosphere_512":
        s = scenes.chain_scene(level=5, res=(512, 512), tex_size=64, seed=1)
        return s.clip, s.indices, s.resolution, None, 0
    if name == "uvsphere_128":
        pos, uv, idx = scenes.uv_sphere(32, 33)
        m = scenes.world_to_clip(np.array([0.3, 0.8, 2.7]), 1.0)
        return scenes.to_clip(pos, m), idx, (128, 128), None, 0
    kind, seed = name.split("_")[0], int(name.split("_")[-1])
    v, idx = soup(seed)
    cull = {"soup": 0, "soupback": 1, "soupfront": 2}[kind]
    return v, idx, (97, 131), None, cull


ALL_CASES = list(FORWARD_CASES) + ["icosphere_512", "uvsphere_128", "soup_1", "soup_2", "soupback_3", "soupfront_4",
                                   "tiny_5", "tiny_6"]


@pytest.mark.parametrize("case", ALL_CASES)
def test_rasterize_forward_bit_exact(rasterizer, case):
    pos, idx, res, vp, cull = _case(case)
    cull_name = ["none", "back", "front"][cull]
    raster, deriv = rasterizer.forward(pos, idx, res, cull_name, True, vp)
    ref_r, ref_d, _ = oracle.rasterize_forward(pos, idx, res, cull, True, vp)
    bits_equal(raster, ref_r)
    bits_equal(deriv, ref_d)
    # without derivatives: same raster, empty derivative array (bindings.cpp:436-447)
    raster2, deriv2 = rasterizer.forward(pos, idx, res, cull_name, False, vp)
    bits_equal(raster2, ref_r)
    assert deriv2.shape == (0, 0, 4)


def test_native_reference_checks(rasterizer):
    """tests/native_tests.cpp:125-157,235-249 and tests/python_smoke.py:18-32,136-143."""
    raster, deriv = rasterizer.forward(TRIANGLE, ONE, (32, 32))
    assert raster[..., 3].sum() > 100
    grad = np.zeros_like(raster)
    grad[..., 0] = (raster[..., 3] > 0).astype(np.float32)
    g = rasterizer.backward(TRIANGLE, ONE, raster, grad)
    assert g.shape == (3, 4) and np.isfinite(g).all()
    eps = 1e-3
    pp, pn = TRIANGLE.copy(), TRIANGLE.copy()
    pp[0, 0] += eps
    pn[0, 0] -= eps
    numeric = (rasterizer.forward(pp, ONE, (32, 32))[0][..., 0].sum() -
               rasterizer.forward(pn, ONE, (32, 32))[0][..., 0].sum()) / (2 * eps)
    assert abs(g[0, 0] - numeric) < 0.2
    vr, _ = rasterizer.forward(TRIANGLE, ONE, (32, 32), viewport=(8.0, 4.0, 16.0, 20.0))
    ys, xs = np.nonzero(vr[..., 3])
    assert xs.min() >= 8 and xs.max() < 24 and ys.min() >= 4 and ys.max() < 24
    sr, _ = rasterizer.forward(TRIANGLE, ONE, (20, 24), viewport=(4.0, 3.0, 16.0, 12.0))
    assert np.count_nonzero(sr[:3, ..., 3]) == 0 and np.count_nonzero(sr[15:, ..., 3]) == 0


@pytest.mark.parametrize("case", ["native_triangle_32", "native_viewport_32", "icosphere_512", "uvsphere_128",
                                  "soup_1", "soupback_3", "tiny_5"])
def test_rasterize_backward(rasterizer, case):
    pos, idx, res, vp, cull = _case(case)
    raster, _, _ = oracle.rasterize_forward(pos, idx, res, cull, False, vp)
    rng = np.random.default_rng(11)
    grad = rng.standard_normal(raster.shape).astype(np.float32)
    grad[rng.random(res) < 0.1, :2] = 0.0
    g = rasterizer.backward(pos, idx, raster, grad, vp)
    ref = oracle.rasterize_backward(pos, idx, raster, grad, vp)
    assert (g[:, 2] == 0).all()
    # per-triangle sums in the reference's pixel order, vertex sums in its corner order
    bits_equal(g, ref)


@pytest.mark.parametrize("variant", ["moved", "random_ids", "other_viewport"])
def test_rasterize_backward_foreign_raster(rasterizer, variant):
    """A raster that the forward did not produce for these positions: pixels may name
    triangles that do not cover them.  The reference's backward loop only compares ids
    inside each triangle's box (rasterize_backward.hlsl:88-100), so those pixels count."""
    pos, idx, res, vp, cull = _case("soup_2")
    rng = np.random.default_rng(21)
    if variant == "moved":
        moved = pos + rng.normal(0, 0.02, pos.shape).astype(np.float32) * pos[:, 3:4]
        raster, _, _ = oracle.rasterize_forward(moved, idx, res, 0, False, vp)
    elif variant == "random_ids":
        raster, _, _ = oracle.rasterize_forward(pos, idx, res, 0, False, vp)
        ids = rng.integers(-2, idx.shape[0] + 3, size=res).astype(np.float32)
        ids[rng.random(res) < 0.2] += 0.5  # half-way ids round to even
        raster[..., 3] = ids
    else:
        raster, _, _ = oracle.rasterize_forward(pos, idx, res, 0, False, (5.0, -3.0, 120.0, 90.0))
    grad = rng.standard_normal(raster.shape).astype(np.float32)
    g = rasterizer.backward(pos, idx, raster, grad, vp)
    ref = oracle.rasterize_backward(pos, idx, raster, grad, vp)
    bits_equal(g, ref)


@pytest.mark.parametrize("attrs", [1, 2, 3, 8, 11])
def test_interpolate(rasterizer, attrs):
    s = scenes.chain_scene(level=4, res=(160, 192), tex_size=32, seed=attrs)
    raster, _, _ = oracle.rasterize_forward(s.clip, s.indices, s.resolution)
    rng = np.random.default_rng(attrs)
    attr = rng.standard_normal((s.clip.shape[0], attrs)).astype(np.float32)
    out = rasterizer.interpolate_forward(attr, s.indices, raster)
    bits_equal(out, oracle.interpolate_forward(attr, s.indices, raster))
    grad = rng.standard_normal(out.shape).astype(np.float32)
    ga, gr = rasterizer.interpolate_backward(attr, s.indices, raster, grad)
    ref_a, ref_r = oracle.interpolate_backward(attr, s.indices, raster, grad)
    bits_equal(gr, ref_r)
    close_sum(ga, ref_a)


@pytest.mark.parametrize("mode", ["clamp", "wrap", "mirror"])
@pytest.mark.parametrize("channels", [1, 3, 4])
def test_texture(rasterizer, mode, channels):
    s = scenes.chain_scene(level=4, res=(150, 170), tex_size=48, channels=channels, seed=channels)
    raster, _, _ = oracle.rasterize_forward(s.clip, s.indices, s.resolution)
    uv = oracle.interpolate_forward(s.uv, s.indices, raster)
    if mode != "clamp":
        uv = (uv * 3.0 - 1.25).astype(np.float32)
    out = rasterizer.texture_forward(s.texture, uv, raster, mode)
    bits_equal(out, oracle.texture_forward(s.texture, uv, raster, mode))
    grad = np.random.default_rng(5).standard_normal(out.shape).astype(np.float32)
    gt, gu = rasterizer.texture_backward(s.texture, uv, raster, grad, mode, True)
    ref_t, ref_u = oracle.texture_backward(s.texture, uv, raster, grad, mode, True)
    bits_equal(gu, ref_u)
    close_sum(gt, ref_t)
    gt2, gu2 = rasterizer.texture_backward(s.texture, uv, raster, grad, mode, False)
    assert gu2 is None
    close_sum(gt2, ref_t)


def test_texture_backward_errors(rasterizer):
    raster, _ = rasterizer.forward(TRIANGLE, ONE, (16, 16))
    tex = np.ones((4, 4, 3), np.float32)
    uv = np.full((16, 16, 2), 0.5, np.float32)
    grad = np.ones((16, 16, 3), np.float32)
    bad = uv.copy()
    covered = np.argwhere(raster[..., 3] > 0)[0]
    bad[tuple(covered)] = np.nan
    with pytest.raises(ValueError, match="uv contains a non-finite value in a covered pixel"):
        rasterizer.texture_backward(tex, bad, raster, grad)
    big = uv.copy()
    big[tuple(covered)] = 1e9
    with pytest.raises(IndexError, match="uv is too large for portable texture addressing"):
        rasterizer.texture_backward(tex, big, raster, grad)


def test_validation_messages(rasterizer):
    with pytest.raises(IndexError, match="triangle_indices contains an invalid vertex index"):
        rasterizer.forward(TRIANGLE, np.array([[0, 1, 3]], np.uint32), (8, 8))
    with pytest.raises(ValueError, match="viewport must contain finite values and positive dimensions"):
        rasterizer.forward(TRIANGLE, ONE, (8, 8), viewport=(0.0, 0.0, 0.0, 4.0))
    with pytest.raises(ValueError, match="cull_mode must be"):
        rasterizer.forward(TRIANGLE, ONE, (8, 8), "sideways")
    with pytest.raises(ValueError, match="positions must have shape"):
        rasterizer.forward(TRIANGLE[:, :3].copy(), ONE, (8, 8))
    raster, _ = rasterizer.forward(TRIANGLE, ONE, (8, 8))
    with pytest.raises(IndexError, match="invalid attribute vertex index"):
        rasterizer.interpolate_forward(np.ones((2, 2), np.float32), ONE, raster)
    with pytest.raises(ValueError, match="address_mode must be"):
        rasterizer.texture_forward(np.ones((2, 2, 1), np.float32), np.zeros((8, 8, 2), np.float32), raster, "tile")


def test_torch_device_path_matches_numpy_path(rasterizer):
    torch = pytest.importorskip("torch")
    import photara_drender_amd as pdr

    s = scenes.chain_scene(level=4, res=(96, 128), tex_size=32, channels=3, seed=3)
    dev = torch.device("cuda:0")
    pos = torch.tensor(s.clip, device=dev, requires_grad=True)
    idx = torch.tensor(s.indices.astype(np.int64), device=dev)
    uvv = torch.tensor(s.uv, device=dev, requires_grad=True)
    tex = torch.tensor(s.texture, device=dev, requires_grad=True)
    raster, deriv = pdr.rasterize(pos, idx, s.resolution, rasterizer=rasterizer)
    uv = pdr.interpolate(uvv, raster, idx, rasterizer=rasterizer)
    img = pdr.texture(tex, uv, raster, rasterizer=rasterizer)
    target = torch.tensor(s.target, device=dev)
    ((img - target) ** 2).sum().backward()
    ref_r, ref_d, _ = oracle.rasterize_forward(s.clip, s.indices, s.resolution)
    bits_equal(raster.detach().cpu().numpy(), ref_r)
    bits_equal(deriv.cpu().numpy(), ref_d)
    ref_uv = oracle.interpolate_forward(s.uv, s.indices, ref_r)
    bits_equal(uv.detach().cpu().numpy(), ref_uv)
    ref_img = oracle.texture_forward(s.texture, ref_uv, ref_r)
    bits_equal(img.detach().cpu().numpy(), ref_img)
    g_img = (2.0 * (ref_img.astype(np.float32) - s.target)).astype(np.float32)
    ref_gt, ref_guv = oracle.texture_backward(s.texture, ref_uv, ref_r, g_img)
    close_sum(tex.grad.cpu().numpy(), ref_gt)
    ref_ga, ref_gr = oracle.interpolate_backward(s.uv, s.indices, ref_r, ref_guv)
    close_sum(uvv.grad.cpu().numpy(), ref_ga, rel=1e-3)
    ref_gp = oracle.rasterize_backward(s.clip, s.indices, ref_r, ref_gr)
    close_sum(pos.grad.cpu().numpy(), ref_gp, rel=1e-3)


def test_batched_views_equal_single_views(rasterizer):
    """pdr_*_batched over uneven views == the single-view entry points, per view."""
    torch = pytest.importorskip("torch")
    import ctypes

    from photara_drender_amd import _native

    sc = scenes.multiview_scene(nx=60, ny=40, views=5, attributes=3, seed=9,
                                sizes=[(64, 64), (48, 80), (100, 70), (33, 17), (128, 96)])
    dev = torch.device("cuda:0")
    V, T, A = sc.clips[0].shape[0], sc.indices.shape[0], 3
    ctx = _native.Context(0)
    lib = _native.lib()
    idx = torch.tensor(sc.indices.view(np.int32), device=dev)
    attr = torch.tensor(sc.attributes, device=dev)
    n = len(sc.clips)
    clips = [torch.tensor(c, device=dev) for c in sc.clips]
    rasters = [torch.empty((h, w, 4), device=dev) for h, w in sc.resolutions]
    derivs = [torch.empty((h, w, 4), device=dev) for h, w in sc.resolutions]
    outs = [torch.empty((h, w, A), device=dev) for h, w in sc.resolutions]
    grads = [torch.tensor(t, device=dev) for t in sc.targets]
    grs = [torch.empty((h, w, 4), device=dev) for h, w in sc.resolutions]
    gpos = [torch.empty((V, 4), device=dev) for _ in range(n)]
    ga = torch.empty((V, A), device=dev)
    rv = (_native.RasterView * n)()
    iv = (_native.InterpView * n)()
    for k, (h, w) in enumerate(sc.resolutions):
        rv[k].positions, rv[k].width, rv[k].height = clips[k].data_ptr(), w, h
        rv[k].viewport = (ctypes.c_float * 4)(0, 0, 0, 0)
        rv[k].raster, rv[k].derivatives = rasters[k].data_ptr(), derivs[k].data_ptr()
        rv[k].grad_raster, rv[k].grad_positions = grs[k].data_ptr(), gpos[k].data_ptr()
        iv[k].width, iv[k].height, iv[k].raster = w, h, rasters[k].data_ptr()
        iv[k].out, iv[k].grad, iv[k].grad_raster = outs[k].data_ptr(), grads[k].data_ptr(), grs[k].data_ptr()
    s = torch.cuda.current_stream().cuda_stream
    _native.check(lib.pdr_rasterize_forward_batched(ctx.handle, rv, n, V, idx.data_ptr(), T, 0, 1, s))
    _native.check(lib.pdr_interpolate_forward_batched(ctx.handle, iv, n, attr.data_ptr(), V, A, idx.data_ptr(), T, s))
    _native.check(lib.pdr_interpolate_backward_batched(ctx.handle, iv, n, attr.data_ptr(), V, A, idx.data_ptr(), T,
                                                       ga.data_ptr(), s))
    _native.check(lib.pdr_rasterize_backward_batched(ctx.handle, rv, n, V, idx.data_ptr(), T, s))
    torch.cuda.synchronize()
    ga_ref = np.zeros((V, A), np.float64)
    for k in range(n):
        ref_r, ref_d, _ = oracle.rasterize_forward(sc.clips[k], sc.indices, sc.resolutions[k])
        bits_equal(rasters[k].cpu().numpy(), ref_r)
        bits_equal(derivs[k].cpu().numpy(), ref_d)
        bits_equal(outs[k].cpu().numpy(), oracle.interpolate_forward(sc.attributes, sc.indices, ref_r))
        g_a, g_r = oracle.interpolate_backward(sc.attributes, sc.indices, ref_r, sc.targets[k])
        bits_equal(grs[k].cpu().numpy(), g_r)
        ga_ref += g_a
        bits_equal(gpos[k].cpu().numpy(), oracle.rasterize_backward(sc.clips[k], sc.indices, ref_r, g_r))
    close_sum(ga.cpu().numpy(), ga_ref)
