"""Pins of the oracle's refinement-precompute functions and of the seam-pair
generator (no GPU), each against what the mathematics or an independent walk
fixes, chosen so that a transposed matrix, a dropped clamp or a wrong edge key
fails:

* ``transform_positions`` (src/texture_refiner.cpp:36-50, clip = M [p, 1] with
  M row-major) against a float64 product with non-symmetric perspective
  look-at matrices;
* ``refine_view_inputs`` (src/texture_refiner.cpp:226-242) in closed form:
  target = first three channels, mask = clamp(visibility, 0, 1) on covered
  pixels and 0 elsewhere (1 without a mask), valid weight = the double sum;
* ``scenes.seam_uv_pairs`` (examples/refine_colmap_texture.py:36-73) against a
  brute-force dictionary walk over every face edge, as a set of sample pairs,
  and on the 1,000 x 500-quad, 256-chart grid against the closed-form count
  (22,500 chart-border edges x 4 samples) and the continuity of the surface
  coordinate across every pair.
"""
import math

import numpy as np
import pytest

import oracle
from photara_drender_amd import scenes


def _cameras(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        eye = scenes.random_direction(rng) * rng.uniform(2.0, 4.0)
        out.append(scenes.world_to_clip(eye, aspect=rng.uniform(0.6, 1.9), fov=rng.uniform(30, 70),
                                        target=rng.uniform(-0.3, 0.3, 3)))
    return out


def test_transform_positions_matches_float64_row_major_product():
    rng = np.random.default_rng(11)
    pos = rng.uniform(-1.5, 1.5, (500, 3)).astype(np.float32)
    for m in _cameras(4, 12):
        assert not np.allclose(m, m.T)  # a transposed read would change the result
        got = oracle.transform_positions(pos, m)
        p = np.concatenate([pos.astype(np.float64), np.ones((pos.shape[0], 1))], 1)
        want = p @ m.astype(np.float64).T  # clip_r = sum_c M[r, c] p_c
        scale = np.abs(p) @ np.abs(m.astype(np.float64)).T
        # four float32 products and three float32 sums: |error| <= 4 u * sum |M_rc p_c|
        assert np.all(np.abs(got - want) <= 4 * 2.0 ** -24 * scale + 1e-30)
        wrong = p @ m.astype(np.float64)  # what a transposed (column-major) read computes
        assert np.abs(wrong - want).max() > 1e-2
    # the scene helper used by the tests agrees with it
    m = _cameras(1, 13)[0]
    np.testing.assert_allclose(oracle.transform_positions(pos, m), scenes.to_clip(pos, m), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("channels", [3, 4])
@pytest.mark.parametrize("with_mask", [False, True])
def test_refine_view_inputs_closed_form(channels, with_mask):
    rng = np.random.default_rng(5 + channels)
    h, w = 13, 17
    image = rng.uniform(0, 1, (h, w, channels)).astype(np.float32)
    raster = np.zeros((h, w, 4), np.float32)
    covered = rng.uniform(size=(h, w)) < 0.6
    raster[..., 3] = np.where(covered, rng.integers(1, 50, (h, w)), 0).astype(np.float32)
    raster[..., :3] = rng.uniform(0, 1, (h, w, 3))  # barycentrics and depth play no part
    vis = None
    if with_mask:
        vis = rng.uniform(-0.5, 1.5, (h, w)).astype(np.float32)  # values outside [0, 1] are clamped
        vis[0, :4] = [-2.0, 0.0, 1.0, 3.0]
    target, mask, weight = oracle.refine_view_inputs(image, raster, vis)
    np.testing.assert_array_equal(target.reshape(h, w, 3), image[..., :3])
    want = np.where(covered, np.clip(vis, 0.0, 1.0) if with_mask else 1.0, 0.0).astype(np.float32)
    np.testing.assert_array_equal(mask.reshape(h, w), want)
    # the valid weight is the double-precision sum of the float masks, in pixel order
    s = 0.0
    for v in want.reshape(-1):
        s += float(v)
    assert weight == s
    assert weight == pytest.approx(math.fsum(want.reshape(-1).astype(np.float64)), rel=1e-15)


def brute_force_pairs(indices, uv, remap, charts, samples):
    """Every face edge (v_a, v_b) into a dictionary keyed by the unordered pair of
    original vertices; keys seen exactly twice on faces of different charts are
    seams, each side sampled from its lower to its higher original vertex."""
    edges = {}
    for f, tri in enumerate(indices):
        for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            oa, ob = int(remap[a]), int(remap[b])
            lo_v, hi_v = (a, b) if oa <= ob else (b, a)
            edges.setdefault((min(oa, ob), max(oa, ob)), []).append((f, lo_v, hi_v))
    t = (np.arange(samples, dtype=np.float32) + np.float32(0.5)) / np.float32(samples)
    out = set()
    for sides in edges.values():
        if len(sides) != 2 or charts[sides[0][0]] == charts[sides[1][0]]:
            continue
        pts = []
        for _, lo_v, hi_v in sides:
            pts.append([tuple((uv[lo_v] * (np.float32(1.0) - tk) + uv[hi_v] * tk).astype(np.float32)) for tk in t])
        for k in range(samples):
            out.add(frozenset([pts[0][k], pts[1][k]]))
    return out


@pytest.mark.parametrize("nx,ny,cx,cy", [(12, 8, 3, 2), (21, 10, 4, 3)])
def test_seam_pairs_equal_brute_force_walk(nx, ny, cx, cy):
    sc = scenes.atlas_scene(nx=nx, ny=ny, charts_x=cx, charts_y=cy, atlas_size=256, gutter=2, seed=9)
    got = scenes.seam_uv_pairs(sc.indices, sc.uv, sc.vertex_remap, sc.face_chart_ids, samples_per_edge=4)
    want = brute_force_pairs(sc.indices, sc.uv, sc.vertex_remap, sc.face_chart_ids, 4)
    got_set = {frozenset([tuple(p[0]), tuple(p[1])]) for p in got}
    assert len(got_set) == got.shape[0]  # no duplicates
    assert got_set == want
    # chart borders: (cx - 1) vertical lines of ny edges, (cy - 1) horizontal lines of nx edges
    assert got.shape[0] == 4 * ((cx - 1) * ny + (cy - 1) * nx)


def test_seam_pairs_skip_edges_inside_a_chart_and_boundary_edges():
    # two triangles of one chart sharing an edge, and a lone triangle: no seams;
    # the same shared edge across two charts: one seam of 3 samples
    uv = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], np.float32)
    idx = np.array([[0, 1, 2], [1, 3, 2]], np.uint32)
    remap = np.arange(4, dtype=np.uint32)
    assert scenes.seam_uv_pairs(idx, uv, remap, np.array([0, 0], np.uint32), 3).shape == (0, 2, 2)
    uv2 = np.array([[0, 0], [0.4, 0], [0, 0.4], [0.6, 0.6], [1, 0.6], [0.6, 1]], np.float32)
    idx2 = np.array([[0, 1, 2], [4, 5, 3]], np.uint32)
    remap2 = np.array([0, 1, 2, 1, 3, 2], np.uint32)  # UV vertices 3 and 5 duplicate geometric 1 and 2
    p = scenes.seam_uv_pairs(idx2, uv2, remap2, np.array([0, 1], np.uint32), 3)
    assert p.shape == (3, 2, 2)
    # side 0 runs from original vertex 1 to 2 along face 0: (0.4, 0) -> (0, 0.4); side 1 (0.6, 0.6) -> (0.6, 1)
    t = (np.arange(3, dtype=np.float32) + np.float32(0.5)) / np.float32(3)
    side0 = uv2[1] * (1 - t[:, None]) + uv2[2] * t[:, None]
    side1 = uv2[3] * (1 - t[:, None]) + uv2[5] * t[:, None]
    got = {frozenset([tuple(a), tuple(b)]) for a, b in p}
    assert got == {frozenset([tuple(a), tuple(b)]) for a, b in zip(side0, side1)}


def test_seam_pairs_on_the_256_chart_grid():
    """The 1,000 x 500-quad grid in 16 x 16 charts: 15 vertical border lines of
    500 edges and 15 horizontal ones of 1,000 edges, 4 samples each = 90,000 pairs,
    and both sides of every pair are the same point of the surface."""
    sc = scenes.atlas_scene()
    p = scenes.seam_uv_pairs(sc.indices, sc.uv, sc.vertex_remap, sc.face_chart_ids)
    assert p.shape == (90000, 2, 2)
    assert 4 * (15 * 500 + 15 * 1000) == 90000
    # invert each chart's affine uv map (atlas_scene: cell (ci, cj) of 128 x 128
    # texels, 2-texel gutter, quads xs / ys of the chart spread over its interior)
    nx, ny, n = 1000, 500, 16
    qx = np.array_split(np.arange(nx), n)
    qy = np.array_split(np.arange(ny), n)
    cell = 2048 / n

    def surface(uvs):
        x, y = uvs[:, 0] * 2048, uvs[:, 1] * 2048
        ci, cj = np.floor(x / cell).astype(int), np.floor(y / cell).astype(int)
        x0 = np.array([q[0] for q in qx])[ci]
        x1 = np.array([q[-1] + 1 for q in qx])[ci]
        y0 = np.array([q[0] for q in qy])[cj]
        y1 = np.array([q[-1] + 1 for q in qy])[cj]
        sx = x0 + (x - (ci * cell + 2)) / (cell - 4) * (x1 - x0)
        sy = y0 + (y - (cj * cell + 2)) / (cell - 4) * (y1 - y0)
        return np.stack([sx / nx, sy / ny], 1), ci + n * cj

    s0, c0 = surface(p[:, 0].astype(np.float64))
    s1, c1 = surface(p[:, 1].astype(np.float64))
    assert np.all(c0 != c1)  # the two sides lie in different charts
    assert np.abs(s0 - s1).max() < 2e-5  # float32 uv rounding only
