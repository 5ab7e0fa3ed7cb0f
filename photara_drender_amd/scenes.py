"""Seeded synthetic scenes for the benchmark workloads and tests.

Every generator is deterministic in its seed (``numpy.random.default_rng``).
Cameras follow the reference convention: a row-major ``world_to_clip`` with
``clip = M @ [x, y, z, 1]`` and GL-style depth in [-1, 1]
(``python/photara_drender/colmap.py:131-153``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


# ---------------------------------------------------------------- meshes


def icosphere(level: int, radius: float = 1.0, jitter: float = 0.0, seed: int = 0):
    """Icosahedron subdivided ``level`` times: 20*4**level triangles.

    Each vertex is scaled by ``radius * (1 + U(-jitter, jitter))``.
    Returns (positions[V,3] f32, indices[T,3] u32).
    """
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache: dict = {}
        nf = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    pos = np.array(verts, np.float64)
    rng = np.random.default_rng(seed)
    scale = radius * (1.0 + rng.uniform(-jitter, jitter, size=(pos.shape[0], 1)))
    return (pos * scale).astype(np.float32), np.array(faces, np.uint32)


def spherical_uv(positions: np.ndarray) -> np.ndarray:
    """u = 0.5 + atan2(z, x)/(2 pi), v = 0.5 + asin(y/|p|)/pi."""
    p = positions.astype(np.float64)
    r = np.linalg.norm(p, axis=1)
    u = 0.5 + np.arctan2(p[:, 2], p[:, 0]) / (2 * math.pi)
    v = 0.5 + np.arcsin(np.clip(p[:, 1] / r, -1, 1)) / math.pi
    return np.stack([u, v], 1).astype(np.float32)


def uv_sphere(segments: int = 32, rings: int = 33, radius: float = 1.0):
    """UV sphere with 2*segments*(rings-1) triangles (poles duplicated per segment)."""
    pos, uv = [], []
    for r in range(rings):
        theta = math.pi * r / (rings - 1)
        for s in range(segments + 1):
            phi = 2 * math.pi * s / segments
            pos.append((radius * math.sin(theta) * math.cos(phi), radius * math.cos(theta),
                        radius * math.sin(theta) * math.sin(phi)))
            uv.append((s / segments, 1.0 - r / (rings - 1)))
    idx = []
    row = segments + 1
    for r in range(rings - 1):
        for s in range(segments):
            a, b = r * row + s, r * row + s + 1
            c, d = (r + 1) * row + s, (r + 1) * row + s + 1
            idx += [(a, c, b), (b, c, d)]
    return np.array(pos, np.float32), np.array(uv, np.float32), np.array(idx, np.uint32)


def heightfield(nx: int = 400, ny: int = 250, extent=(1.0, 0.625), sigma: float = 0.05, seed: int = 2):
    """Grid of nx*ny quads over [-ex,ex]x[-ey,ey] (2*nx*ny triangles), z ~ N(0, sigma)."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(-extent[0], extent[0], nx + 1)
    ys = np.linspace(-extent[1], extent[1], ny + 1)
    gx, gy = np.meshgrid(xs, ys)
    gz = rng.normal(0.0, sigma, size=gx.shape)
    pos = np.stack([gx, gz, gy], -1).reshape(-1, 3).astype(np.float32)  # y up
    i = np.arange(ny)[:, None] * (nx + 1) + np.arange(nx)[None, :]
    a, b, c, d = i, i + 1, i + nx + 1, i + nx + 2
    idx = np.stack([np.stack([a, c, b], -1), np.stack([b, c, d], -1)], 2).reshape(-1, 3).astype(np.uint32)
    return pos, idx


# ---------------------------------------------------------------- cameras


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> np.ndarray:
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    if np.linalg.norm(s) < 1e-9:
        s = np.cross(f, (1.0, 0.0, 0.0))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m


def perspective(fov_y_deg: float, aspect: float, near: float = 0.1, far: float = 10.0) -> np.ndarray:
    f = 1.0 / math.tan(math.radians(fov_y_deg) / 2)
    m = np.zeros((4, 4))
    m[0, 0], m[1, 1] = f / aspect, f
    m[2, 2], m[2, 3] = (far + near) / (near - far), 2 * far * near / (near - far)
    m[3, 2] = -1.0
    return m


def world_to_clip(eye, aspect, fov=45.0, target=(0.0, 0.0, 0.0), near=0.1, far=10.0) -> np.ndarray:
    """Row-major float32 [4,4]."""
    return (perspective(fov, aspect, near, far) @ look_at(eye, target)).astype(np.float32)


def random_direction(rng) -> np.ndarray:
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def to_clip(positions: np.ndarray, m: np.ndarray) -> np.ndarray:
    """float32 clip positions [V,4] = M @ [p,1] (host-side helper for tests and scenes)."""
    p = np.concatenate([positions.astype(np.float64), np.ones((positions.shape[0], 1))], 1)
    return (p @ m.astype(np.float64).T).astype(np.float32)


# ---------------------------------------------------------------- workloads


@dataclass
class ChainScene:
    """One view of a textured mesh: rasterize -> interpolate(uv) -> texture."""

    clip: np.ndarray  # [V,4] f32
    indices: np.ndarray  # [T,3] u32
    uv: np.ndarray  # [V,2] f32
    texture: np.ndarray  # [th,tw,C] f32
    target: np.ndarray  # [H,W,C] f32
    resolution: tuple  # (H, W)


def chain_scene(level=5, res=(512, 512), tex_size=1024, channels=4, seed=1) -> ChainScene:
    """Plain fwd+bwd chain: jittered icosphere, spherical UV, U[0,1] texture and target,
    look-at camera from a random direction at distance 3, fov 45 deg."""
    rng = np.random.default_rng(seed)
    pos, idx = icosphere(level, 1.0, 0.01, seed)
    uv = spherical_uv(pos)
    eye = 3.0 * random_direction(rng)
    m = world_to_clip(eye, res[1] / res[0], 45.0)
    tex = rng.uniform(0, 1, size=(tex_size, tex_size, channels)).astype(np.float32)
    target = rng.uniform(0, 1, size=(res[0], res[1], channels)).astype(np.float32)
    return ChainScene(to_clip(pos, m), idx, uv, tex, target, res)


MULTIVIEW_SIZES = [(512, 512), (480, 640), (600, 800), (768, 1024), (720, 1280), (1024, 1280), (900, 1600),
                   (1080, 1920)]  # (H, W), used twice


@dataclass
class MultiViewScene:
    clips: list  # per view [V,4] f32
    indices: np.ndarray
    attributes: np.ndarray  # [V,A]
    targets: list  # per view [H,W,A]
    resolutions: list  # (H,W)


def multiview_scene(nx=400, ny=250, views=16, attributes=8, seed=2, sizes=None) -> MultiViewScene:
    """Batched multi-view raster: heightfield (2*nx*ny tris), N(0,1) attributes and
    targets, ``views`` cameras looking at the grid from 30-80 deg elevation."""
    rng = np.random.default_rng(seed)
    pos, idx = heightfield(nx, ny, seed=seed)
    attrs = rng.normal(size=(pos.shape[0], attributes)).astype(np.float32)
    sizes = sizes or (MULTIVIEW_SIZES * ((views + len(MULTIVIEW_SIZES) - 1) // len(MULTIVIEW_SIZES)))[:views]
    clips, targets = [], []
    for k, (h, w) in enumerate(sizes):
        elev = math.radians(rng.uniform(30, 80))
        azim = rng.uniform(0, 2 * math.pi)
        dist = rng.uniform(2.2, 2.8)
        eye = dist * np.array([math.cos(elev) * math.cos(azim), math.sin(elev), math.cos(elev) * math.sin(azim)])
        m = world_to_clip(eye, w / h, 45.0)
        clips.append(to_clip(pos, m))
        targets.append(rng.normal(size=(h, w, attributes)).astype(np.float32))
    return MultiViewScene(clips, idx, attrs, targets, list(sizes))


# ---------------------------------------------------------------- atlas refinement scene


@dataclass
class AtlasScene:
    positions: np.ndarray  # [V_uv, 3] f32 (geometry duplicated at chart borders)
    uv: np.ndarray  # [V_uv, 2] f32 atlas coordinates
    indices: np.ndarray  # [T, 3] u32
    vertex_remap: np.ndarray  # [V_uv] u32 index of the shared geometric vertex
    face_chart_ids: np.ndarray  # [T] u32
    surface: np.ndarray  # [V_uv, 2] f32 continuous surface coordinates in [0,1]^2
    chart_texels: np.ndarray  # [H, W] i32 chart id per texel (-1 in gutters)
    atlas_size: int


def atlas_scene(nx=1000, ny=500, charts_x=16, charts_y=16, atlas_size=2048, gutter=2, sigma=0.02, seed=3,
                extent=(1.0, 0.5)) -> AtlasScene:
    """Perturbed grid of nx*ny quads (2*nx*ny triangles), cut into charts_x*charts_y
    rectangular charts; chart (i,j) maps to the interior of atlas cell (i,j) with a
    ``gutter``-texel margin.  UV vertices are duplicated on chart borders."""
    rng = np.random.default_rng(seed)
    gx = np.linspace(-extent[0], extent[0], nx + 1)
    gy = np.linspace(-extent[1], extent[1], ny + 1)
    heights = rng.normal(0.0, sigma, size=(ny + 1, nx + 1))
    qx = np.array_split(np.arange(nx), charts_x)
    qy = np.array_split(np.arange(ny), charts_y)
    cell_w, cell_h = atlas_size / charts_x, atlas_size / charts_y
    pos, uv, surf, remap, idx, fchart = [], [], [], [], [], []
    chart_tex = np.full((atlas_size, atlas_size), -1, np.int32)
    base = 0
    for cj, ys in enumerate(qy):
        for ci, xs in enumerate(qx):
            x0, x1, y0, y1 = xs[0], xs[-1] + 1, ys[0], ys[-1] + 1
            vx = np.arange(x0, x1 + 1)
            vy = np.arange(y0, y1 + 1)
            mx, my = np.meshgrid(vx, vy)
            n_vx, n_vy = vx.size, vy.size
            pos.append(np.stack([gx[mx], heights[my, mx], gy[my]], -1).reshape(-1, 3))
            surf.append(np.stack([mx / nx, my / ny], -1).reshape(-1, 2))
            remap.append((my * (nx + 1) + mx).reshape(-1))
            # chart interior in texels: [cx0+gutter, cx1-gutter] mapped linearly over the chart's quads
            tx0 = ci * cell_w + gutter
            tx1 = (ci + 1) * cell_w - gutter
            ty0 = cj * cell_h + gutter
            ty1 = (cj + 1) * cell_h - gutter
            u = (tx0 + (mx - x0) / max(x1 - x0, 1) * (tx1 - tx0)) / atlas_size
            v = (ty0 + (my - y0) / max(y1 - y0, 1) * (ty1 - ty0)) / atlas_size
            uv.append(np.stack([u, v], -1).reshape(-1, 2))
            a = base + (np.arange(n_vy - 1)[:, None] * n_vx + np.arange(n_vx - 1)[None, :])
            b, c, d = a + 1, a + n_vx, a + n_vx + 1
            tri = np.stack([np.stack([a, c, b], -1), np.stack([b, c, d], -1)], 2).reshape(-1, 3)
            idx.append(tri)
            fchart.append(np.full(tri.shape[0], cj * charts_x + ci))
            chart_tex[int(np.floor(ty0)):int(np.ceil(ty1)), int(np.floor(tx0)):int(np.ceil(tx1))] = cj * charts_x + ci
            base += n_vx * n_vy
    return AtlasScene(np.concatenate(pos).astype(np.float32), np.concatenate(uv).astype(np.float32),
                      np.concatenate(idx).astype(np.uint32), np.concatenate(remap).astype(np.uint32),
                      np.concatenate(fchart).astype(np.uint32), np.concatenate(surf).astype(np.float32), chart_tex,
                      atlas_size)


def seam_uv_pairs(indices, uv, vertex_remap, face_chart_ids, samples_per_edge=4) -> np.ndarray:
    """Seam sample pairs (examples/refine_colmap_texture.py:36-73): edges keyed by
    their sorted original-vertex pair; keys used by exactly two faces of different
    charts; ``samples_per_edge`` points at (k+0.5)/n along both sides.
    Returns [N, 2, 2] f32 (pair order is free: losses and gradients do not depend on it)."""
    keys, lo_uv, hi_uv, charts = [], [], [], []
    for first, second in ((0, 1), (1, 2), (2, 0)):
        va, vb = indices[:, first], indices[:, second]
        oa, ob = vertex_remap[va].astype(np.uint64), vertex_remap[vb].astype(np.uint64)
        swap = oa > ob
        keys.append((np.minimum(oa, ob) << np.uint64(32)) | np.maximum(oa, ob))
        lo_uv.append(np.where(swap[:, None], uv[vb], uv[va]))
        hi_uv.append(np.where(swap[:, None], uv[va], uv[vb]))
        charts.append(face_chart_ids)
    key = np.concatenate(keys)
    lo, hi, ch = np.concatenate(lo_uv), np.concatenate(hi_uv), np.concatenate(charts)
    order = np.argsort(key, kind="stable")
    key, lo, hi, ch = key[order], lo[order], hi[order], ch[order]
    starts = np.r_[0, np.flatnonzero(key[1:] != key[:-1]) + 1]
    ends = np.r_[starts[1:], key.size]
    starts = starts[(ends - starts) == 2]
    starts = starts[ch[starts] != ch[starts + 1]]
    if starts.size == 0:
        return np.empty((0, 2, 2), np.float32)
    t = ((np.arange(samples_per_edge, dtype=np.float32) + 0.5) / samples_per_edge)[None, :, None]
    first = lo[starts, None] * (1.0 - t) + hi[starts, None] * t
    second = lo[starts + 1, None] * (1.0 - t) + hi[starts + 1, None] * t
    return np.ascontiguousarray(np.stack((first, second), axis=2).reshape(-1, 2, 2), dtype=np.float32)


def ground_truth_params(seed: int = 3):
    rng = np.random.default_rng(seed + 100)
    return [(rng.uniform(2, 24, size=2), rng.uniform(0, 2 * math.pi), rng.uniform(0.03, 0.1, size=3))
            for _ in range(6)]


def ground_truth(surface, seed: int = 3):
    """Procedural RGB in surface coordinates: 6 seeded sinusoids + a checker, clamped to [0,1].
    ``surface`` [..., 2] is a numpy array (float32 result) or a torch tensor (same device)."""
    if hasattr(surface, "data_ptr"):
        import torch

        s = surface.to(torch.float64)
        out = torch.full(s.shape[:-1] + (3,), 0.5, dtype=torch.float64, device=s.device)
        for f, ph, amp in ground_truth_params(seed):
            wave = torch.sin(2 * math.pi * (float(f[0]) * s[..., 0:1] + float(f[1]) * s[..., 1:2]) + ph)
            out += torch.tensor(amp, dtype=torch.float64, device=s.device) * wave
        checker = torch.remainder(torch.floor(s[..., 0] * 16) + torch.floor(s[..., 1] * 8), 2)[..., None]
        return torch.clamp(out + 0.15 * (checker - 0.5), 0.0, 1.0).to(torch.float32)
    s = surface.astype(np.float64)
    out = np.full(s.shape[:-1] + (3,), 0.5)
    for f, ph, amp in ground_truth_params(seed):
        out += amp * np.sin(2 * math.pi * (f[0] * s[..., 0:1] + f[1] * s[..., 1:2]) + ph)
    checker = ((np.floor(s[..., 0] * 16) + np.floor(s[..., 1] * 8)) % 2)[..., None]
    out += 0.15 * (checker - 0.5)
    return np.clip(out, 0.0, 1.0).astype(np.float32)


def atlas_views(scene: AtlasScene, views=64, size=(768, 1024), seed=3, fov=45.0):
    """world_to_clip matrices of cameras above the grid (elevation 35-75 deg)."""
    rng = np.random.default_rng(seed + 200)
    mats = []
    h, w = size
    for _ in range(views):
        elev = math.radians(rng.uniform(35, 75))
        azim = rng.uniform(0, 2 * math.pi)
        dist = rng.uniform(1.0, 1.8)
        target = np.array([rng.uniform(-0.6, 0.6), 0.0, rng.uniform(-0.3, 0.3)])
        eye = target + dist * np.array([math.cos(elev) * math.cos(azim), math.sin(elev), math.cos(elev) * math.sin(azim)])
        mats.append(world_to_clip(eye, w / h, fov, target=target, near=0.05, far=10.0))
    return np.stack(mats).astype(np.float32)


def camera_center(world_to_clip_matrix) -> np.ndarray:
    """World position of a perspective camera: the point whose clip x, y and w vanish."""
    m = np.asarray(world_to_clip_matrix, np.float64).reshape(4, 4)
    rows = m[[0, 1, 3]]
    return np.linalg.solve(rows[:, :3], -rows[:, 3]).astype(np.float32)


def vertex_normals(positions, indices) -> np.ndarray:
    """Area-weighted unit vertex normals of a triangle mesh."""
    p = np.asarray(positions, np.float64)
    t = np.asarray(indices, np.int64)
    fn = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    n = np.zeros_like(p)
    for c in range(3):
        np.add.at(n, t[:, c], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n.astype(np.float32)


def ellipse_masks(views, size, seed=3, count=3):
    """Seeded binary visibility masks: 1 except inside a few random ellipses."""
    rng = np.random.default_rng(seed + 300)
    h, w = size
    yy, xx = np.mgrid[0:h, 0:w]
    masks = []
    for _ in range(views):
        m = np.ones((h, w), np.float32)
        for _ in range(count):
            cx, cy = rng.uniform(0, w), rng.uniform(0, h)
            ax, ay = rng.uniform(0.04, 0.15) * w, rng.uniform(0.04, 0.15) * h
            m[((xx - cx) / ax) ** 2 + ((yy - cy) / ay) ** 2 < 1] = 0.0
        masks.append(m)
    return masks
