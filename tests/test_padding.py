"""pad_texture_atlas (src/texture_padding.cpp:226-351).

CPU, pinning the oracle:
* the reference's known answers - tests/native_tests.cpp:34-117 (24x24 atlas,
  margin 3 with the unobserved fill: 16 valid, 64 unobserved filled, guard and
  untouched texels as asserted; disabled padding fills nothing and reports 64
  remaining) and the padding half of tests/texture_surface_fill_test.cpp:45-57
  (earlier fills seed gutters and are counted; repeating the padding changes
  nothing);
* the 8SSEDT field against brute-force nearest sources: a single source fills
  exactly the disc of radius `margin`; with many sources the chosen source is
  the nearest for at least 99 % of texels and never more than one texel
  farther;
* the unobserved fill honours maximum_fill_distance and fill_unobserved.
GPU, through the C ABI: colour bits, filled mask and statistics equal the
oracle on random masks of many shapes (odd widths, single rows and columns,
atlases wider than 8192 and 12000 texels), also when the parallel band
sweep gives up waiting and the ordered fallback computes the field, and the
package's numpy entry point behaves as python/photara_drender/baking.py:33-80.
"""
import numpy as np
import pytest

import oracle


def native_case():
    S = 24
    col = np.zeros((S, S, 4), np.float32)
    val = np.zeros((S, S), np.float32)
    cov = np.zeros((S, S), np.float32)
    col[2:6, 2:6] = [0.25, 0.5, 0.75, 1.0]
    val[2:6, 2:6] = 1
    cov[2:6, 2:6] = 1
    cov[10:18, 10:18] = 1
    return col, val, cov


def test_reference_known_answers_native_tests():
    col, val, cov = native_case()
    c, f, st = oracle.pad_texture_atlas(col, val, 3, cov, None, fill_unobserved=True)
    assert st["valid_texels"] == 16 and st["unobserved_texels"] == 64 and st["remaining_unobserved"] == 0
    assert st["gutter_texels"] > 0 and st["filled_texels"] == st["gutter_texels"] + st["unobserved_texels"]
    np.testing.assert_array_equal(c[6, 2], [0.25, 0.5, 0.75, 1.0])  # (x=2, y=6) inside the guard band
    np.testing.assert_array_equal(c[10, 10], [0.25, 0.5, 0.75, 1.0])
    assert f[10, 10] == 1 and f[2, 2] == 0
    assert c[12, 2, 0] == 0  # outside guard band and coverage
    _, _, st = oracle.pad_texture_atlas(np.zeros_like(col), val, 0, cov, None, fill_unobserved=False)
    assert st["filled_texels"] == 0 and st["remaining_unobserved"] == 64


def surface_fill_case():
    # the atlas of tests/texture_surface_fill_test.cpp after fill_texture_on_surface
    col = np.zeros((1, 7, 4), np.float32)
    col[0, 0] = [0.8, 0, 0, 1]
    col[0, 2] = [0, 0, 0.6, 1]
    col[0, 1] = [0.8, 0, 0, 1]
    col[0, 3] = [0.8, 0, 0, 1]
    val = np.array([[1, 0, 1, 0, 0, 0, 0]], np.float32)
    cov = np.array([[1, 1, 1, 1, 1, 0, 0]], np.float32)
    filled = np.array([[0, 1, 0, 1, 0, 0, 0]], np.float32)
    return col, val, cov, filled


def test_reference_known_answers_surface_fill_padding():
    col, val, cov, filled = surface_fill_case()
    c, f, st = oracle.pad_texture_atlas(col, val, 2, cov, filled, fill_unobserved=False)
    assert st["unobserved_texels"] == 2 and st["remaining_unobserved"] == 1
    assert c[0, 5, 0] == np.float32(0.8) and c[0, 4, 0] == 0
    assert f[0, 3] == 1
    c2, f2, st2 = oracle.pad_texture_atlas(c, val, 2, cov, f, fill_unobserved=False)
    assert st2["gutter_texels"] == st["gutter_texels"] and st2["unobserved_texels"] == st["unobserved_texels"]


def test_single_source_fills_the_disc():
    H, W, m = 41, 37, 9
    col = np.zeros((H, W, 3), np.float32)
    val = np.zeros((H, W), np.float32)
    col[17, 20] = [0.3, 0.6, 0.9]
    val[17, 20] = 1
    c, f, st = oracle.pad_texture_atlas(col, val, m, None, None)
    yy, xx = np.mgrid[0:H, 0:W]
    disc = ((yy - 17) ** 2 + (xx - 20) ** 2 <= m * m) & (val == 0)
    np.testing.assert_array_equal(f == 1, disc)
    assert (c[disc] == col[17, 20]).all() and st["gutter_texels"] == disc.sum()


def test_field_is_the_nearest_source():
    rng = np.random.default_rng(4)
    H, W = 48, 64
    val = (rng.uniform(size=(H, W)) < 0.02).astype(np.float32)
    ids = np.cumsum(val.ravel()).reshape(H, W) * val
    col = np.stack([ids, np.zeros_like(ids), np.zeros_like(ids)], -1).astype(np.float32)
    c, f, _ = oracle.pad_texture_atlas(col, val, 1000, None, None)
    src = np.argwhere(val > 0)
    yy, xx = np.mgrid[0:H, 0:W]
    d2 = ((yy[..., None] - src[:, 0]) ** 2 + (xx[..., None] - src[:, 1]) ** 2)
    best = np.sqrt(d2.min(-1))
    chosen = np.sqrt(d2[yy, xx, c[..., 0].astype(np.int64) - 1])
    empty = val == 0
    assert (chosen[empty] == best[empty]).mean() >= 0.99
    assert (chosen - best)[empty].max() <= 1.0


def test_unobserved_fill_options():
    col, val, cov = native_case()
    _, _, st = oracle.pad_texture_atlas(col, val, 3, cov, None, fill_unobserved=True, maximum_fill_distance=8)
    # chart texels (10..17)^2 nearer than 8 texels to the block (2..5)^2: distance from (5,5)
    yy, xx = np.mgrid[10:18, 10:18]
    near = ((yy - 5) ** 2 + (xx - 5) ** 2 <= 64).sum()
    assert st["unobserved_texels"] == near and st["remaining_unobserved"] == 64 - near
    with pytest.raises(ValueError, match="3 or 4 channels"):
        oracle.pad_texture_atlas(col[..., :2], val, 3)


# ---------------------------------------------------------------------- GPU


def _case(rng, H, W, C, density):
    col = rng.uniform(0, 1, (H, W, C)).astype(np.float32)
    seeds = rng.uniform(size=(H, W)) < density
    val = np.zeros((H, W), np.float32)
    val[seeds] = 1.0
    val[rng.uniform(size=(H, W)) < 0.02] = 0.4  # below the threshold
    cov = (rng.uniform(size=(H, W)) < 0.5).astype(np.float32)
    filled = ((rng.uniform(size=(H, W)) < 0.05) & (val < 0.5)).astype(np.float32)
    col[val < 0.5] = 0
    return col, val, cov, filled


def _gpu_pad(ctx, col, val, cov, filled, margin, fill_unobserved, max_fill):
    import ctypes

    import torch

    from photara_drender_amd import _native

    c = torch.tensor(col, device="cuda")
    v = torch.tensor(val, device="cuda")
    cv = None if cov is None else torch.tensor(cov, device="cuda")
    f = None if filled is None else torch.tensor(filled, device="cuda")
    o = _native.PaddingOptions(margin, int(fill_unobserved), max_fill, 0.5)
    s = _native.PaddingStats()
    _native.check(_native.lib().pdr_pad_texture_atlas(
        ctx.handle, c.data_ptr(), col.shape[1], col.shape[0], col.shape[2], v.data_ptr(),
        None if cv is None else cv.data_ptr(), None if f is None else f.data_ptr(), ctypes.byref(o),
        ctypes.byref(s), torch.cuda.current_stream().cuda_stream))
    stats = {"valid_texels": s.valid_texels, "gutter_texels": s.gutter_texels,
             "unobserved_texels": s.unobserved_texels, "remaining_unobserved": s.remaining_unobserved,
             "filled_texels": s.filled_texels}
    return c.cpu().numpy(), None if f is None else f.cpu().numpy(), stats


SHAPES = [(1, 1), (1, 37), (37, 1), (48, 64), (131, 97), (64, 256), (300, 301), (8, 12288), (4, 16384), (1000, 3),
          (517, 2100)]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
def test_gpu_padding_matches_oracle(shape):
    from photara_drender_amd import _native

    ctx = _native.Context(0, flags=0)
    rng = np.random.default_rng(shape[0] * 7919 + shape[1])
    H, W = shape
    for k, (margin, fill, cap, C, with_cov, with_filled, density) in enumerate([
            (8, True, 0, 4, True, True, 0.01), (2, False, 0, 3, True, False, 0.05), (0, True, 5, 4, True, True, 0.002),
            (3, False, 0, 3, False, False, 0.03), (8, True, 0, 3, False, True, 0.0)]):
        col, val, cov, filled = _case(rng, H, W, C, density)
        cov = cov if with_cov else None
        filled = filled if with_filled else None
        rc, rf, rs = oracle.pad_texture_atlas(col, val, margin, cov, filled, fill, cap)
        gc, gf, gs = _gpu_pad(ctx, col, val, cov, filled, margin, fill, cap)
        np.testing.assert_array_equal(gc.view(np.uint32), rc.view(np.uint32), err_msg=f"case {k}")
        if filled is not None:
            np.testing.assert_array_equal(gf, rf, err_msg=f"case {k}")
        assert gs == rs, (k, gs, rs)


@pytest.mark.gpu
def test_gpu_padding_falls_back_when_a_band_gives_up_waiting():
    """The parallel band sweep's bounded wait forced to give up (a poll limit of
    1 set from the host through a test hook; every kernel still exits): the
    field is recomputed one band per launch, in order, and colour bits, filled
    mask and statistics equal the oracle's and the parallel path's."""
    from photara_drender_amd import _native

    lib = _native.lib()
    ctx = _native.Context(0, flags=0)
    rng = np.random.default_rng(77)
    H, W = 517, 300  # nine bands of 64 rows
    col, val, cov, filled = _case(rng, H, W, 4, 0.01)
    rc, rf, rs = oracle.pad_texture_atlas(col, val, 6, cov, filled, True, 0)
    pc, pf, ps = _gpu_pad(ctx, col, val, cov, filled, 6, True, 0)
    before = lib.pdr_testing_pad_fallbacks()
    lib.pdr_testing_set_pad_poll_limit(1)
    try:
        fc, ff, fs = _gpu_pad(ctx, col, val, cov, filled, 6, True, 0)
    finally:
        lib.pdr_testing_set_pad_poll_limit(0)
    assert lib.pdr_testing_pad_fallbacks() == before + 1  # the wait gave up and the ordered sweep ran
    for c, f, st in ((pc, pf, ps), (fc, ff, fs)):
        np.testing.assert_array_equal(c.view(np.uint32), rc.view(np.uint32))
        np.testing.assert_array_equal(f, rf)
        assert st == rs
    # the default bound again: no fallback
    _gpu_pad(ctx, col, val, cov, filled, 6, True, 0)
    assert lib.pdr_testing_pad_fallbacks() == before + 1


@pytest.mark.gpu
def test_gpu_padding_of_a_baked_atlas_and_numpy_entry_point():
    import photara_drender_amd as pdr
    from photara_drender_amd import scenes

    sc = scenes.atlas_scene(nx=60, ny=30, charts_x=4, charts_y=4, atlas_size=256, gutter=2, seed=8)
    mats = scenes.atlas_views(sc, 6, (96, 128), seed=8)
    rng = np.random.default_rng(8)
    imgs = [rng.uniform(0, 1, (96, 128, 3)).astype(np.float32) for _ in range(6)]
    cams = np.stack([scenes.camera_center(m) for m in mats])
    nrm = scenes.vertex_normals(sc.positions, sc.indices)
    color, conf, src, valid, cov, face = oracle.bake(sc.positions, nrm, sc.uv, sc.indices, imgs, mats, cams,
                                                     resolution=(256, 256))
    padded = pdr.pad_texture_atlas(color, valid, 8)
    ref, _, _ = oracle.pad_texture_atlas(color, valid, 8, None, None, fill_unobserved=False)
    np.testing.assert_array_equal(padded.view(np.uint32), ref.view(np.uint32))
    assert np.count_nonzero(padded[..., :3]) >= np.count_nonzero(color[..., :3])  # tests/python_smoke.py:101-103
    with pytest.raises(ValueError, match="valid_mask must match"):
        pdr.pad_texture_atlas(color, valid[:, :5], 8)
