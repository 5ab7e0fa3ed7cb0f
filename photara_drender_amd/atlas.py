"""Textured rendering, the atlas losses and a PyTorch Adam loop over atlas texels:
the Python helpers of python/photara_drender/atlas.py:14-257 (names, arguments,
defaults, validation messages and results), on this package's zero-copy
autograd functions.

The losses define the refinement objective (python/photara_drender/atlas.py:
65-144), with rho(d) = sqrt(d^2 + eps^2):
* masked photometric: sum(rho(rendered - target) * m) / (max(sum m, 1) * C);
* chart-aware total variation: rho of the forward differences in x and y of
  texel pairs that are both valid and in one chart, summed and divided by
  (max(pairs, 1) * C);
* seam consistency: the mean of rho over the channels of the colour difference
  between the two bilinear atlas samples of every seam pair.

optimize_texture_atlas runs torch.optim.Adam on the atlas with one view per
step in round-robin order (step s uses view s mod views); each step's loss is
the photometric term, then + tv_weight * TV, + seam_weight * seam, +
prior_weight * mean(valid * (atlas - initial)^2), added in that order; after
each step the atlas is clamped and the loss appended to the history.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, NamedTuple, Optional, Sequence, Tuple

import torch

from ._photara_drender import Rasterizer
from .torch import interpolate, rasterize
from .torch import texture as _sample


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


def _rho(d: torch.Tensor, epsilon: float) -> torch.Tensor:
    """Charbonnier penalty sqrt(d^2 + eps^2), elementwise."""
    return torch.sqrt(d.square() + epsilon * epsilon)


def _coverage(raster: torch.Tensor, dtype) -> torch.Tensor:
    """[H, W, 1] 1 where the raster names a triangle."""
    return (raster[..., 3:4] > 0).to(dtype=dtype)


def render_textured_mesh(clip_positions, indices, vertex_uv, texture_values, resolution, *, rasterizer: Rasterizer,
                         viewport=None, cull_mode: str = "none", address_mode: str = "clamp") -> TexturedRender:
    """Rasterize (with barycentric derivatives), interpolate the vertex uv, sample the
    texture; the intermediates are returned for the losses."""
    raster, derivatives = rasterize(clip_positions, indices, resolution, rasterizer=rasterizer, viewport=viewport,
                                    cull_mode=cull_mode, output_barycentric_derivatives=True)
    pixel_uv = interpolate(vertex_uv, raster, indices, rasterizer=rasterizer)
    image = _sample(texture_values, pixel_uv, raster, rasterizer=rasterizer, address_mode=address_mode)
    return TexturedRender(image, _coverage(raster, image.dtype), raster, derivatives, pixel_uv)


def masked_charbonnier_loss(rendered, target, valid_mask, *, epsilon: float = 1e-3):
    """Robust photometric loss normalised by the number of valid pixels and channels."""
    if rendered.shape != target.shape:
        raise ValueError("rendered and target must have the same shape")
    if valid_mask.shape != rendered.shape[:-1] + (1,):
        raise ValueError("valid_mask must have shape [height, width, 1]")
    pixels = valid_mask.sum().clamp_min(1.0)
    return (_rho(rendered - target, epsilon) * valid_mask).sum() / (pixels * rendered.shape[-1])


def _neighbour_pairs(valid: torch.Tensor, chart_ids: Optional[torch.Tensor]):
    """Masks of the horizontal ([H, W-1]) and vertical ([H-1, W]) texel pairs that
    the total variation counts: both texels valid and, with charts, in one chart."""
    pairs = []
    for axis in (1, 0):
        n = valid.shape[axis] - 1
        first, second = valid.narrow(axis, 0, n), valid.narrow(axis, 1, n)
        keep = first & second
        if chart_ids is not None:
            keep = keep & (chart_ids.narrow(axis, 0, n) == chart_ids.narrow(axis, 1, n))
        pairs.append(keep)
    return pairs


def atlas_total_variation(texture_values, *, valid_mask=None, chart_ids=None, epsilon: float = 1e-4):
    """Charbonnier total variation over neighbouring texels of one chart (no blur across seams)."""
    if texture_values.ndim != 3:
        raise ValueError("texture_values must have shape [height, width, channels]")
    size = tuple(texture_values.shape[:2])
    device = texture_values.device
    if valid_mask is None:
        valid = torch.ones(size, dtype=torch.bool, device=device)
    else:
        valid = valid_mask.to(device=device, dtype=torch.bool)
    if tuple(valid.shape) != size:
        raise ValueError("valid_mask must have shape [texture_height, texture_width]")
    if chart_ids is not None and tuple(chart_ids.shape) != size:
        raise ValueError("chart_ids must have shape [texture_height, texture_width]")
    charts = None if chart_ids is None else chart_ids.to(device=device)
    along_x, along_y = _neighbour_pairs(valid, charts)
    dx = texture_values[:, 1:] - texture_values[:, :-1]
    dy = texture_values[1:, :] - texture_values[:-1, :]
    total = (_rho(dx, epsilon) * along_x[..., None]).sum() + (_rho(dy, epsilon) * along_y[..., None]).sum()
    count = along_x.sum() + along_y.sum()
    return total / (count.clamp_min(1) * texture_values.shape[-1])


_COVERED_ROWS: dict = {}


def _covered_row(n: int, device) -> torch.Tensor:
    """A 1 x n raster naming triangle 1 at every pixel: the texture op then samples
    all n uv points (texture_forward.hlsl samples only covered pixels)."""
    key = (n, str(device))
    row = _COVERED_ROWS.get(key)
    if row is None:
        row = torch.zeros((1, n, 4), dtype=torch.float32, device=device)
        row[..., 3] = 1.0
        _COVERED_ROWS.clear()  # keep one: the seam set of a run does not change
        _COVERED_ROWS[key] = row
    return row


def _sample_points(texture_values, uv_points, *, rasterizer: Rasterizer, address_mode: str) -> torch.Tensor:
    """Bilinear atlas colours [N, C] at uv points [N, 2] (differentiable in the atlas)."""
    n = uv_points.shape[0]
    out = _sample(texture_values, uv_points.reshape(1, n, 2).contiguous(), _covered_row(n, texture_values.device),
                  rasterizer=rasterizer, address_mode=address_mode)
    return out.reshape(n, texture_values.shape[-1])


def seam_consistency_loss(texture_values, seam_uv_pairs, *, rasterizer: Rasterizer, address_mode: str = "clamp",
                          epsilon: float = 1e-4):
    """Penalise colour disagreement between the two uv samples of each seam pair."""
    if seam_uv_pairs.ndim != 3 or tuple(seam_uv_pairs.shape[1:]) != (2, 2):
        raise ValueError("seam_uv_pairs must have shape [seam_count, 2, 2]")
    count = seam_uv_pairs.shape[0]
    if count == 0:
        return texture_values.sum() * 0.0  # zero, still connected to the atlas
    colours = _sample_points(texture_values, seam_uv_pairs.reshape(2 * count, 2), rasterizer=rasterizer,
                             address_mode=address_mode).reshape(count, 2, -1)
    return _rho(colours[:, 0] - colours[:, 1], epsilon).mean()


class _Views:
    """The views of the loop: per view its target, its mask (coverage times the
    optional external visibility mask) and its atlas sample.  Without caching
    each step renders the view; with ``cache_view_rasters`` the raster and pixel
    uv of a view are kept after its first step and the atlas is re-sampled."""

    def __init__(self, clips, indices, vertex_uv, targets, viewports, masks, rasterizer, options):
        self.clips, self.indices, self.vertex_uv = clips, indices, vertex_uv
        self.targets, self.viewports, self.masks = targets, viewports, masks
        self.rasterizer, self.options = rasterizer, options
        self.kept: dict = {}

    def __len__(self) -> int:
        return self.clips.shape[0]

    def _external(self, k: int, like: torch.Tensor) -> torch.Tensor:
        m = self.masks[k]
        if m is None:
            return like
        m = m[..., None] if m.ndim == 2 else m
        return like * m.to(device=like.device, dtype=like.dtype)

    def sample(self, k: int, atlas: torch.Tensor):
        """(rendered image, valid mask) of view k with the current atlas."""
        mode = self.options.address_mode
        kept = self.kept.get(k)
        if kept is not None:
            raster, pixel_uv, covered = kept
            image = _sample(atlas, pixel_uv, raster, rasterizer=self.rasterizer, address_mode=mode)
        else:
            r = render_textured_mesh(self.clips[k], self.indices, self.vertex_uv, atlas, self.targets[k].shape[:2],
                                     rasterizer=self.rasterizer, viewport=self.viewports[k], address_mode=mode)
            image, covered = r.image, r.valid_mask
            if self.options.cache_view_rasters:
                self.kept[k] = (r.raster.detach(), r.uv.detach(), r.valid_mask.detach())
        return image, self._external(k, covered)


def _regularisers(initial: torch.Tensor, atlas_valid_mask, chart_ids, seam_uv_pairs, rasterizer,
                  options: AtlasOptimizationOptions) -> List[Callable[[torch.Tensor], torch.Tensor]]:
    """The weighted atlas-only terms of the loss, in the order they are added."""
    terms = []
    if options.total_variation_weight != 0:
        terms.append(lambda a: options.total_variation_weight *
                     atlas_total_variation(a, valid_mask=atlas_valid_mask, chart_ids=chart_ids))
    if options.seam_weight != 0 and seam_uv_pairs is not None:
        terms.append(lambda a: options.seam_weight *
                     seam_consistency_loss(a, seam_uv_pairs, rasterizer=rasterizer, address_mode=options.address_mode))
    if options.initial_prior_weight != 0:
        def prior(a):
            sq = (a - initial).square()
            if atlas_valid_mask is not None:
                sq = sq * atlas_valid_mask.to(device=sq.device, dtype=sq.dtype)[..., None]
            return options.initial_prior_weight * sq.mean()

        terms.append(prior)
    return terms


def optimize_texture_atlas(initial_texture, clip_positions_by_view, indices, vertex_uv, target_images: Sequence, *,
                           rasterizer: Rasterizer, viewports=None, visibility_masks=None, atlas_valid_mask=None,
                           chart_ids=None, seam_uv_pairs=None,
                           options: AtlasOptimizationOptions = AtlasOptimizationOptions()
                           ) -> Tuple[torch.Tensor, List[float]]:
    """Optimise the atlas texels against several calibrated views with
    torch.optim.Adam, one view per step in round-robin order.

    ``clip_positions_by_view`` is [view_count, vertex_count, 4].  Returns the
    optimised atlas (detached) and the per-step losses."""
    if clip_positions_by_view.ndim != 3 or clip_positions_by_view.shape[2] != 4:
        raise ValueError("clip_positions_by_view must have shape [view_count, vertex_count, 4]")
    count = clip_positions_by_view.shape[0]
    if count == 0 or len(target_images) != count:
        raise ValueError("target_images must contain one image per view")
    viewports = [None] * count if viewports is None else viewports
    visibility_masks = [None] * count if visibility_masks is None else visibility_masks
    if len(viewports) != count or len(visibility_masks) != count:
        raise ValueError("viewports and visibility_masks must match the view count")
    if options.steps <= 0 or options.learning_rate <= 0:
        raise ValueError("optimization steps and learning_rate must be positive")

    views = _Views(clip_positions_by_view, indices, vertex_uv, target_images, viewports, visibility_masks, rasterizer,
                   options)
    atlas = torch.nn.Parameter(initial_texture.detach().clone())
    extra_terms = _regularisers(initial_texture.detach().clone(), atlas_valid_mask, chart_ids, seam_uv_pairs,
                                rasterizer, options)
    adam = torch.optim.Adam([atlas], lr=options.learning_rate)
    history: List[float] = []
    for step in range(options.steps):
        adam.zero_grad(set_to_none=True)
        k = step % len(views)
        image, valid = views.sample(k, atlas)
        loss = masked_charbonnier_loss(image, target_images[k], valid, epsilon=options.photometric_epsilon)
        for term in extra_terms:
            loss = loss + term(atlas)
        loss.backward()
        adam.step()
        with torch.no_grad():
            atlas.clamp_(options.clamp_min, options.clamp_max)
        history.append(float(loss.detach().cpu()))
    return atlas.detach(), history
