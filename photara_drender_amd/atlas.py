"""Textured rendering, atlas losses and the PyTorch Adam loop
(python/photara_drender/atlas.py:14-257), over this package's autograd functions.

The loss definitions follow the reference:
* masked Charbonnier: sum(sqrt(d^2 + eps^2) * m) / (max(sum m, 1) * C)   (atlas.py:65-81)
* chart-aware total variation over forward differences in x and y whose two
  texels are valid and in the same chart, divided by (max(pairs, 1) * C)  (atlas.py:84-115)
* seam consistency: mean Charbonnier (eps 1e-4) of the colour difference at
  the two uv samples of each seam pair                                   (atlas.py:118-144)
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, NamedTuple, Optional, Sequence, Tuple

import torch

from ._photara_drender import Rasterizer
from .torch import interpolate, rasterize
from .torch import texture as sample_texture


class TexturedRender(NamedTuple):
    image: torch.Tensor
    valid_mask: torch.Tensor
    raster: torch.Tensor
    barycentric_derivatives: torch.Tensor
    uv: torch.Tensor


@dataclass(frozen=True)
class AtlasOptimizationOptions:
    steps: int = 150
    learning_rate: float = 1e-2
    photometric_epsilon: float = 1e-3
    total_variation_weight: float = 1e-4
    seam_weight: float = 1e-3
    initial_prior_weight: float = 0.0
    address_mode: str = "clamp"
    clamp_min: float = 0.0
    clamp_max: float = 1.0
    cache_view_rasters: bool = False


def render_textured_mesh(clip_positions, indices, vertex_uv, texture_values, resolution, *, rasterizer: Rasterizer,
                         viewport=None, cull_mode: str = "none", address_mode: str = "clamp") -> TexturedRender:
    """rasterize -> interpolate(uv) -> texture, keeping the intermediates the losses need."""
    raster, derivatives = rasterize(clip_positions, indices, resolution, rasterizer=rasterizer, viewport=viewport,
                                    cull_mode=cull_mode, output_barycentric_derivatives=True)
    uv = interpolate(vertex_uv, raster, indices, rasterizer=rasterizer)
    image = sample_texture(texture_values, uv, raster, rasterizer=rasterizer, address_mode=address_mode)
    covered = (raster[..., 3:4] > 0).to(dtype=image.dtype)
    return TexturedRender(image, covered, raster, derivatives, uv)


def masked_charbonnier_loss(rendered, target, valid_mask, *, epsilon: float = 1e-3):
    """Robust photometric loss normalised by the number of valid pixels."""
    if rendered.shape != target.shape:
        raise ValueError("rendered and target must have the same shape")
    if valid_mask.shape != rendered.shape[:-1] + (1,):
        raise ValueError("valid_mask must have shape [height, width, 1]")
    robust = ((rendered - target) ** 2 + epsilon ** 2).sqrt()
    count = valid_mask.sum().clamp_min(1.0) * rendered.shape[-1]
    return (robust * valid_mask).sum() / count


def atlas_total_variation(texture_values, *, valid_mask=None, chart_ids=None, epsilon: float = 1e-4):
    """Charbonnier total variation over neighbour pairs inside one chart."""
    if texture_values.ndim != 3:
        raise ValueError("texture_values must have shape [height, width, channels]")
    h, w = texture_values.shape[:2]
    dev = texture_values.device
    valid = torch.ones((h, w), dtype=torch.bool, device=dev) if valid_mask is None else \
        valid_mask.to(device=dev, dtype=torch.bool)
    if valid.shape != (h, w):
        raise ValueError("valid_mask must have shape [texture_height, texture_width]")
    if chart_ids is not None and chart_ids.shape != (h, w):
        raise ValueError("chart_ids must have shape [texture_height, texture_width]")
    pair_x = valid[:, :-1] & valid[:, 1:]
    pair_y = valid[:-1, :] & valid[1:, :]
    if chart_ids is not None:
        c = chart_ids.to(device=dev)
        pair_x = pair_x & (c[:, :-1] == c[:, 1:])
        pair_y = pair_y & (c[:-1, :] == c[1:, :])
    dx = ((texture_values[:, 1:] - texture_values[:, :-1]) ** 2 + epsilon ** 2).sqrt()
    dy = ((texture_values[1:, :] - texture_values[:-1, :]) ** 2 + epsilon ** 2).sqrt()
    total = (dx * pair_x[..., None]).sum() + (dy * pair_y[..., None]).sum()
    pairs = pair_x.sum() + pair_y.sum()
    return total / (pairs.clamp_min(1) * texture_values.shape[-1])


def seam_consistency_loss(texture_values, seam_uv_pairs, *, rasterizer: Rasterizer, address_mode: str = "clamp",
                          epsilon: float = 1e-4):
    """Charbonnier disagreement of the two atlas samples of every seam pair."""
    if seam_uv_pairs.ndim != 3 or tuple(seam_uv_pairs.shape[1:]) != (2, 2):
        raise ValueError("seam_uv_pairs must have shape [seam_count, 2, 2]")
    n = seam_uv_pairs.shape[0]
    if n == 0:
        return texture_values.sum() * 0.0
    uv = seam_uv_pairs.reshape(1, 2 * n, 2).contiguous()
    covered = torch.zeros((1, 2 * n, 4), dtype=torch.float32, device=texture_values.device)
    covered[..., 3] = 1.0
    colours = sample_texture(texture_values, uv, covered, rasterizer=rasterizer, address_mode=address_mode)
    colours = colours.reshape(n, 2, texture_values.shape[-1])
    d = colours[:, 0] - colours[:, 1]
    return ((d ** 2 + epsilon ** 2).sqrt()).mean()


def optimize_texture_atlas(initial_texture, clip_positions_by_view, indices, vertex_uv, target_images, *,
                           rasterizer: Rasterizer, viewports=None, visibility_masks=None, atlas_valid_mask=None,
                           chart_ids=None, seam_uv_pairs=None,
                           options: AtlasOptimizationOptions = AtlasOptimizationOptions()
                           ) -> Tuple[torch.Tensor, List[float]]:
    """torch.optim.Adam over the atlas, one view per step in round-robin order."""
    if clip_positions_by_view.ndim != 3 or clip_positions_by_view.shape[2] != 4:
        raise ValueError("clip_positions_by_view must have shape [view_count, vertex_count, 4]")
    nviews = clip_positions_by_view.shape[0]
    if len(target_images) != nviews or nviews == 0:
        raise ValueError("target_images must contain one image per view")
    viewports = [None] * nviews if viewports is None else viewports
    visibility_masks = [None] * nviews if visibility_masks is None else visibility_masks
    if len(viewports) != nviews or len(visibility_masks) != nviews:
        raise ValueError("viewports and visibility_masks must match the view count")
    if options.steps <= 0 or options.learning_rate <= 0:
        raise ValueError("optimization steps and learning_rate must be positive")
    param = torch.nn.Parameter(initial_texture.detach().clone())
    prior = initial_texture.detach().clone()
    optimizer = torch.optim.Adam([param], lr=options.learning_rate)
    history: List[float] = []
    cache: dict = {}
    for step in range(options.steps):
        optimizer.zero_grad(set_to_none=True)
        k = step % nviews
        target = target_images[k]
        if k in cache:
            raster, uv, valid = cache[k]
            image = sample_texture(param, uv, raster, rasterizer=rasterizer, address_mode=options.address_mode)
        else:
            r = render_textured_mesh(clip_positions_by_view[k], indices, vertex_uv, param, target.shape[:2],
                                     rasterizer=rasterizer, viewport=viewports[k], address_mode=options.address_mode)
            image, valid = r.image, r.valid_mask
            if options.cache_view_rasters:
                cache[k] = (r.raster.detach(), r.uv.detach(), r.valid_mask.detach())
        mask = visibility_masks[k]
        if mask is not None:
            mask = mask[..., None] if mask.ndim == 2 else mask
            valid = valid * mask.to(device=valid.device, dtype=valid.dtype)
        loss = masked_charbonnier_loss(image, target, valid, epsilon=options.photometric_epsilon)
        if options.total_variation_weight != 0:
            loss = loss + options.total_variation_weight * atlas_total_variation(param, valid_mask=atlas_valid_mask,
                                                                                 chart_ids=chart_ids)
        if options.seam_weight != 0 and seam_uv_pairs is not None:
            loss = loss + options.seam_weight * seam_consistency_loss(param, seam_uv_pairs, rasterizer=rasterizer,
                                                                      address_mode=options.address_mode)
        if options.initial_prior_weight != 0:
            diff = (param - prior) ** 2
            if atlas_valid_mask is not None:
                diff = diff * atlas_valid_mask.to(device=diff.device, dtype=diff.dtype)[..., None]
            loss = loss + options.initial_prior_weight * diff.mean()
        loss.backward()
        optimizer.step()
        with torch.no_grad():
            param.clamp_(options.clamp_min, options.clamp_max)
        history.append(float(loss.detach().cpu()))
    return param.detach(), history
