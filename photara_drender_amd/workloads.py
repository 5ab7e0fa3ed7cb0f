"""GPU-resident benchmark problems built from the seeded scenes.

``refine_problem`` builds the atlas-refinement workload: a perturbed grid mesh cut
into rectangular UV charts, cameras above it, targets rendered from a procedural
ground-truth texture (noise + binary masks), seam pairs and the per-view uv maps
used to score an atlas by PSNR.  Everything is rendered with this package's own
kernels and stays on the GPU.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _native, scenes


@dataclass
class RefineProblem:
    positions: object  # [V,3] device f32
    uv: object  # [V,2]
    indices: object  # [T,3] int32 (uint32 bits)
    images: list  # per view [H,W,3] noisy targets
    clean: list  # per view [H,W,3] noise-free renders of the ground truth
    masks: list  # per view [H,W] binary
    world_to_clip: np.ndarray  # [N,4,4] f32
    seam_pairs: object  # [P,2,2]
    chart_texels: object  # [S,S] int32, device-resident like the views
    view_uv: list  # per view [H,W,2] interpolated atlas uv
    view_raster: list  # per view [H,W,4]
    atlas_size: int
    ground_truth_atlas: object  # [S,S,3] ground truth sampled at texel centres (reference point)
    normals: object = None  # [V,3] device f32 area-weighted vertex normals
    camera_positions: np.ndarray = None  # [N,3] f32 camera centres


def _clip(pos, m):
    import torch

    mt = torch.tensor(m, dtype=torch.float32, device=pos.device)
    return (pos @ mt[:, :3].T + mt[:, 3]).contiguous()


def _raster(ctx, clip, idx, h, w):
    import torch

    r = torch.empty((h, w, 4), dtype=torch.float32, device=clip.device)
    _native.check(_native.lib().pdr_rasterize_forward(ctx.handle, clip.data_ptr(), clip.shape[0], idx.data_ptr(),
                                                      idx.shape[0], w, h, 0, 0, None, r.data_ptr(), None, None,
                                                      torch.cuda.current_stream().cuda_stream))
    return r


def _interp(ctx, attr, idx, r):
    import torch

    h, w = r.shape[:2]
    out = torch.empty((h, w, attr.shape[1]), dtype=torch.float32, device=attr.device)
    _native.check(_native.lib().pdr_interpolate_forward(ctx.handle, attr.data_ptr(), attr.shape[0], attr.shape[1],
                                                        idx.data_ptr(), idx.shape[0], r.data_ptr(), w, h,
                                                        out.data_ptr(), torch.cuda.current_stream().cuda_stream))
    return out


def refine_problem(ctx, nx=1000, ny=500, charts=16, atlas_size=2048, views=64, size=(768, 1024), noise=0.02,
                   seed=3, device="cuda:0") -> RefineProblem:
    import torch

    dev = torch.device(device)
    sc = scenes.atlas_scene(nx=nx, ny=ny, charts_x=charts, charts_y=charts, atlas_size=atlas_size, seed=seed)
    mats = scenes.atlas_views(sc, views, size, seed=seed)
    masks_np = scenes.ellipse_masks(views, size, seed=seed)
    pos = torch.tensor(sc.positions, device=dev)
    uv = torch.tensor(sc.uv, device=dev)
    surf = torch.tensor(sc.surface, device=dev)
    idx = torch.tensor(sc.indices.view(np.int32), device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    images, clean, view_uv, view_raster = [], [], [], []
    h, w = size
    for k in range(views):
        r = _raster(ctx, _clip(pos, mats[k]), idx, h, w)
        gt = scenes.ground_truth(_interp(ctx, surf, idx, r), seed)
        covered = (r[..., 3:4] != 0).to(torch.float32)
        gt = gt * covered
        n = torch.randn(gt.shape, generator=gen, device=dev) * noise
        images.append(torch.clamp(gt + n, 0.0, 1.0).contiguous())
        clean.append(gt.contiguous())
        view_uv.append(_interp(ctx, uv, idx, r))
        view_raster.append(r)
    pairs = scenes.seam_uv_pairs(sc.indices, sc.uv, sc.vertex_remap, sc.face_chart_ids)
    # ground-truth colour at every texel centre of the charts (surface coordinates by inverse chart map)
    s = atlas_size
    tex_uv = (np.stack(np.meshgrid(np.arange(s), np.arange(s)), -1) + 0.5) / s
    gt_atlas = np.zeros((s, s, 3), np.float32)
    cell = s / charts
    gutter = 2
    qx = np.array_split(np.arange(nx), charts)
    qy = np.array_split(np.arange(ny), charts)
    for cj, ys in enumerate(qy):
        for ci, xs in enumerate(qx):
            tx0, tx1 = ci * cell + gutter, (ci + 1) * cell - gutter
            ty0, ty1 = cj * cell + gutter, (cj + 1) * cell - gutter
            m = sc.chart_texels == cj * charts + ci
            tu, tv = tex_uv[m][:, 0] * s, tex_uv[m][:, 1] * s
            su = (xs[0] + (tu - tx0) / (tx1 - tx0) * (xs[-1] + 1 - xs[0])) / nx
            sv = (ys[0] + (tv - ty0) / (ty1 - ty0) * (ys[-1] + 1 - ys[0])) / ny
            gt_atlas[m] = scenes.ground_truth(np.stack([su, sv], -1), seed)
    torch.cuda.synchronize()
    normals = torch.tensor(scenes.vertex_normals(sc.positions, sc.indices), device=dev)
    cams = np.stack([scenes.camera_center(m) for m in mats])
    return RefineProblem(pos, uv, idx, images, clean, [torch.tensor(m, device=dev) for m in masks_np], mats,
                         torch.tensor(pairs, device=dev), torch.tensor(sc.chart_texels, device=dev), view_uv,
                         view_raster, atlas_size,
                         torch.tensor(gt_atlas, device=dev), normals, cams)


def bake_initial_atlas(ctx, problem: RefineProblem, blend_mode="weighted_average", padding=8):
    """The plain method's starting atlas: TextureBaker::bake of the noisy views
    (shadow-map visibility, the views' masks), RGB of the baked colour, then
    pad_texture_atlas with a ``padding``-texel guard band as the reference's
    refinement example does before refining (examples/refine_colmap_texture.py:
    240-242).  Returns (atlas [S,S,3], bake result, seconds)."""
    import time

    import torch

    from .bake import bake_atlas, pad_atlas

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = bake_atlas(ctx, problem.positions, problem.normals, problem.uv, problem.indices, problem.images,
                     problem.world_to_clip, problem.camera_positions, problem.masks,
                     resolution=(problem.atlas_size, problem.atlas_size), blend_mode=blend_mode)
    atlas = res.color[..., :3].contiguous()
    if padding:
        pad_atlas(ctx, atlas, res.valid, padding)
    torch.cuda.synchronize()
    return atlas, res, time.perf_counter() - t0


def atlas_psnr(ctx, problem: RefineProblem, atlas, against="clean") -> float:
    """Masked PSNR (MAX=1) of the atlas rendered into every view, over covered pixels
    with mask > 0, against the noise-free renders (or the noisy targets)."""
    import torch

    atlas = atlas.contiguous()
    sse = torch.zeros((), dtype=torch.float64, device=atlas.device)
    count = torch.zeros((), dtype=torch.float64, device=atlas.device)
    s = problem.atlas_size
    refs = problem.clean if against == "clean" else problem.images
    for k in range(len(refs)):
        r, uvk = problem.view_raster[k], problem.view_uv[k]
        h, w = r.shape[:2]
        out = torch.empty((h, w, 3), dtype=torch.float32, device=atlas.device)
        _native.check(_native.lib().pdr_texture_forward(ctx.handle, atlas.data_ptr(), s, s, 3, 0, uvk.data_ptr(),
                                                        r.data_ptr(), w, h, out.data_ptr(),
                                                        torch.cuda.current_stream().cuda_stream))
        valid = ((r[..., 3] != 0) & (problem.masks[k] > 0)).to(torch.float64)
        sse += (((out - refs[k]).to(torch.float64) ** 2).sum(-1) * valid).sum()
        count += valid.sum() * 3
    mse = float(sse / count)
    return 10.0 * math.log10(1.0 / max(mse, 1e-20))


class TexturedChain:
    """The textured raster chain over device tensors: rasterize -> interpolate
    (uv) -> texture (clamp), the loss 0.5 * sum (color - target)^2's gradient (one op), then
    texture, interpolate and rasterize backward to texels, uv attributes and
    clip positions.  ``step(stream)`` runs it eagerly on a stream;
    ``capture(stream)`` records one step into a HIP graph (torch.cuda.CUDAGraph)
    after the caller has run it eagerly on that stream (the context's scratch
    is sized and its launch parameters chosen by then)."""

    NAMES = ("raster", "deriv", "uvs", "color", "g_color", "g_tex", "g_uvs", "g_uv", "g_raster", "g_pos")

    def __init__(self, ctx, scene, device="cuda:0"):
        import torch

        dev = torch.device(device)
        self.ctx = ctx
        self.H, self.W = scene.resolution
        self.th, self.tw, self.C = scene.texture.shape
        self.V, self.T = scene.clip.shape[0], scene.indices.shape[0]
        self.clip = torch.tensor(scene.clip, device=dev)
        self.idx = torch.tensor(scene.indices.view(np.int32), device=dev)
        self.uv = torch.tensor(scene.uv, device=dev)
        self.tex = torch.tensor(scene.texture, device=dev)
        self.target = torch.tensor(scene.target, device=dev)
        H, W, C = self.H, self.W, self.C
        shapes = {"raster": (H, W, 4), "deriv": (H, W, 4), "uvs": (H, W, 2), "color": (H, W, C),
                  "g_color": (H, W, C), "g_tex": (self.th, self.tw, C), "g_uvs": (H, W, 2), "g_uv": (self.V, 2),
                  "g_raster": (H, W, 4), "g_pos": (self.V, 4)}
        for k, s in shapes.items():
            setattr(self, k, torch.empty(s, device=dev))
        self._vp = (ctypes.c_float * 4)()

    def step(self, stream) -> None:
        import torch

        lib, ck, h = _native.lib(), _native.check, self.ctx.handle
        V, T, W, H, C, tw, th = self.V, self.T, self.W, self.H, self.C, self.tw, self.th
        st = stream.cuda_stream
        ck(lib.pdr_rasterize_forward(h, self.clip.data_ptr(), V, self.idx.data_ptr(), T, W, H, 0, 1, None,
                                     self.raster.data_ptr(), self.deriv.data_ptr(), self._vp, st))
        ck(lib.pdr_interpolate_forward(h, self.uv.data_ptr(), V, 2, self.idx.data_ptr(), T, self.raster.data_ptr(),
                                       W, H, self.uvs.data_ptr(), st))
        ck(lib.pdr_texture_forward(h, self.tex.data_ptr(), tw, th, C, 0, self.uvs.data_ptr(), self.raster.data_ptr(),
                                   W, H, self.color.data_ptr(), st))
        with torch.cuda.stream(stream):
            torch.sub(self.color, self.target, alpha=1.0, out=self.g_color)
        ck(lib.pdr_texture_backward(h, self.tex.data_ptr(), tw, th, C, 0, self.uvs.data_ptr(), self.raster.data_ptr(),
                                    W, H, self.g_color.data_ptr(), 1, self.g_tex.data_ptr(), self.g_uvs.data_ptr(),
                                    st))
        ck(lib.pdr_interpolate_backward(h, self.uv.data_ptr(), V, 2, self.idx.data_ptr(), T, self.raster.data_ptr(),
                                        W, H, self.g_uvs.data_ptr(), self.g_uv.data_ptr(), self.g_raster.data_ptr(),
                                        st))
        ck(lib.pdr_rasterize_backward(h, self.clip.data_ptr(), V, self.idx.data_ptr(), T, self.raster.data_ptr(), W,
                                      H, self.g_raster.data_ptr(), None, self.g_pos.data_ptr(), st))

    def capture(self, stream):
        import torch

        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            self.step(stream)
        return graph

    def outputs(self) -> dict:
        return {k: getattr(self, k).cpu().numpy() for k in self.NAMES}
