"""CPU checks of the raster's candidate regions (photara_drender_amd/csrc/raster_geom.hpp,
through libpdr's host-only hook pdr_testing_candidate_region, which runs the same
compiled setup code as the kernels).

The forward tests only the pixels of a triangle's candidate region, and the
backward walks only that region inside the reference's bounding box, so every
pixel that the reference's float coverage test (shaders/rasterize_forward.hlsl:
103-113, epsilon-inclusive) accepts must lie in it.  Coverage comes from the
oracle's single-triangle raster; triangles include slivers, sub-pixel and
sub-epsilon triangles, edges through pixel centres, near-degenerate and
partly off-screen triangles, flipped windings and custom viewports.  The
bound is also checked to stay tight (spans close to the covered area).
"""
import ctypes

import numpy as np
import pytest

import oracle

H, W = 40, 56


def region(clip, viewport, cull=0, allow_cols=True, h=H, w=W):
    """(has candidates, rectangle (x0, x1, y0, y1), [h, w] candidate mask)."""
    from photara_drender_amd import _native

    clip = np.ascontiguousarray(clip, np.float32).reshape(12)
    vp = np.ascontiguousarray(viewport, np.float32)
    rect = np.zeros(4, np.int32)
    mask = np.zeros((h, w), np.uint8)
    ok = _native.lib().pdr_testing_candidate_region(clip.ctypes.data, w, h, vp.ctypes.data, cull, int(allow_cols),
                                                    rect.ctypes.data, mask.ctypes.data)
    return bool(ok), rect, mask.astype(bool)


def covered(clip, viewport, cull=0, h=H, w=W):
    r, _, _ = oracle.rasterize_forward(np.asarray(clip, np.float32).reshape(3, 4), np.array([[0, 1, 2]], np.uint32),
                                       (h, w), cull, False, viewport)
    return r[..., 3] == 1.0


def triangles(rng, n):
    """Clip-space triangles of many shapes (w in [0.5, 2])."""
    out = []
    px = lambda v: (v + 0.5) * 2.0 / np.array([W, H]) - 1.0  # pixel centre -> NDC
    for k in range(n):
        kind = k % 8
        if kind == 0:  # random
            s = rng.uniform(-1.3, 1.3, (3, 2))
        elif kind == 1:  # sliver: two vertices close together
            a = rng.uniform(-1, 1, 2)
            b = rng.uniform(-1, 1, 2)
            s = np.stack([a, b, b + rng.normal(0, 0.01, 2)])
        elif kind == 2:  # sub-pixel
            c = rng.uniform(-1, 1, 2)
            s = c + rng.normal(0, 0.01, (3, 2))
        elif kind == 3:  # sub-epsilon (area ~1e-12..1e-8 in NDC^2)
            c = px(rng.integers(0, [W, H]).astype(np.float64))
            s = c + rng.normal(0, 10.0 ** rng.uniform(-6, -3), (3, 2))
        elif kind == 4:  # vertices on pixel centres (edges through sample points)
            s = px(rng.integers(-3, [W + 3, H + 3], (3, 2)).astype(np.float64))
        elif kind == 5:  # near-degenerate: almost collinear
            a, b = rng.uniform(-1.2, 1.2, (2, 2))
            t = rng.uniform(0, 1)
            s = np.stack([a, b, a + t * (b - a) + rng.normal(0, 1e-4, 2)])
        elif kind == 6:  # large, partly off-screen
            s = rng.uniform(-6, 6, (3, 2))
        else:  # thin and long along a pixel row or column
            a = rng.uniform(-1, 1, 2)
            d = np.array([rng.uniform(-2, 2), rng.normal(0, 0.02)])
            if rng.uniform() < 0.5:
                d = d[::-1]
            s = np.stack([a, a + d, a + d + rng.normal(0, 0.01, 2)])
        w = rng.uniform(0.5, 2.0, 3)
        z = rng.uniform(-0.9, 0.9, 3)
        clip = np.stack([s[:, 0] * w, s[:, 1] * w, z * w, w], 1)
        if rng.uniform() < 0.5:
            clip = clip[::-1]
        out.append(clip.astype(np.float32))
    return out


@pytest.mark.parametrize("allow_cols", [True, False])
@pytest.mark.parametrize("viewport", [(0, 0, W, H), (3.5, -2.25, 47.0, 41.5), (-10, 5, 80, 20)])
def test_candidate_region_holds_every_covered_pixel(viewport, allow_cols):
    """Forward regions (rows or columns) and backward regions (rows) hold every covered pixel."""
    rng = np.random.default_rng(11)
    vp = np.array(viewport, np.float32)
    n_cov = n_reg = n_rect = n_edge = 0
    for clip in triangles(rng, 800):
        cov = covered(clip, vp)
        ok, rect, mask = region(clip, vp, 0, allow_cols)
        if not ok:
            assert not cov.any()
            continue
        x0, x1, y0, y1 = rect
        ys, xs = np.nonzero(cov)
        assert np.all((xs >= x0) & (xs <= x1) & (ys >= y0) & (ys <= y1))
        assert not (cov & ~mask).any(), (clip.tolist(), viewport)
        n_cov += cov.sum()
        n_reg += mask.sum()
        n_rect += (x1 - x0 + 1) * (y1 - y0 + 1)
        # covered pixels whose region neighbour along x or y is outside the region: the bound is tight there
        inner = mask.copy()
        inner[:, 1:] &= mask[:, :-1]
        inner[:, :-1] &= mask[:, 1:]
        inner[1:, :] &= mask[:-1, :]
        inner[:-1, :] &= mask[1:, :]
        n_edge += (cov & ~inner).sum()
    assert n_cov > 1000 and n_edge > 1000
    # regions stay close to coverage (about a pixel of halo per line), well under the rectangles
    assert n_reg < 0.6 * n_rect


def test_candidate_region_respects_cull_modes():
    rng = np.random.default_rng(12)
    vp = np.array((0, 0, W, H), np.float32)
    for clip in triangles(rng, 200):
        for cull in (1, 2):
            ok, _, mask = region(clip, vp, cull)
            cov = covered(clip, vp, cull)
            if cov.any():
                assert ok
                assert not (cov & ~mask).any()


def test_candidate_region_tight_on_slivers():
    """A one-pixel-wide vertical sliver: rows of a few pixels (backward), or one or two
    columns (forward), not the bin's 16 columns."""
    vp = np.array((0, 0, W, H), np.float32)
    x = (20.5 + 0.5) * 2.0 / W - 1.0
    clip = np.array([[x, -0.9, 0.0, 1.0], [x + 0.04, -0.9, 0.0, 1.0], [x + 0.02, 0.9, 0.0, 1.0]], np.float32)
    cov = covered(clip, vp)
    assert cov.sum() > 10
    for allow_cols in (False, True):
        ok, rect, mask = region(clip, vp, 0, allow_cols)
        assert ok and not (cov & ~mask).any()
        assert mask.sum(1).max() <= 3
        assert mask.sum() <= 2 * cov.sum() + 2 * (rect[3] - rect[2] + 1)
