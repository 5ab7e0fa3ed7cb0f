"""Texture bake for arrays already on the GPU (pdr_bake, TextureBaker::bake of
src/texture_baker.cpp:156-361).  ``TextureBaker.bake`` of the native module is
the reference's numpy entry point; this wrapper takes torch tensors (or numpy)
and returns torch tensors on the device of the inputs without host staging.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _native

BLEND_MODES = {"best_view": 0, "weighted_average": 1, "softmax": 2}
VISIBILITY_MODES = {"shadow_map": 0, "hybrid_ray_query": 1, "ray_query": 2}


@dataclass
class BakeResult:
    color: object        # [H, W, 4] (RGB, valid)
    confidence: object   # [H, W]
    source_view: object  # [H, W] int32 bits of the u32 view index (0xFFFFFFFF: none)
    valid: object        # [H, W]
    coverage: object     # [H, W]
    face_ids: object     # [H, W] int32 bits of the u32 face index (0xFFFFFFFF: outside)


def _prepare(context, positions, normals, uv, indices, images, world_to_clip, camera_positions, visibility_masks,
             viewports, weights, depth_bias, min_view_cosine):
    """Device tensors of the mesh and views plus the pdr_bake_view records pointing at them."""
    import torch

    handle = context.handle if hasattr(context, "handle") else context.context_handle
    dev = positions.device if hasattr(positions, "device") else torch.device("cuda:0")

    def t(a, dtype=torch.float32):
        if a is None:
            return None
        if hasattr(a, "data_ptr"):
            return a.detach().to(device=dev, dtype=dtype).contiguous()
        return torch.as_tensor(np.ascontiguousarray(a), device=dev).to(dtype).contiguous()

    pos, nrm, uvs = t(positions), t(normals), t(uv)
    idx = t(indices, torch.int32)
    n = len(images)
    imgs = [t(im) for im in images]
    masks = [None] * n if visibility_masks is None else [t(m) for m in visibility_masks]
    mats = np.asarray(world_to_clip, np.float32).reshape(n, 16)
    cams = np.asarray(camera_positions, np.float32).reshape(n, 3)
    views = (_native.BakeView * n)()
    for k in range(n):
        v = views[k]
        _native.lib().pdr_bake_view_default(ctypes.byref(v))
        v.height, v.width, v.channels = int(imgs[k].shape[0]), int(imgs[k].shape[1]), int(imgs[k].shape[2])
        v.image = imgs[k].data_ptr()
        v.visibility_mask = None if masks[k] is None else masks[k].data_ptr()
        v.world_to_clip = (ctypes.c_float * 16)(*mats[k].tolist())
        v.camera_position = (ctypes.c_float * 3)(*cams[k].tolist())
        vp = None if viewports is None else viewports[k]
        v.has_viewport = 0 if vp is None else 1
        if vp is not None:
            v.viewport = (ctypes.c_float * 4)(*[float(x) for x in vp])
        v.weight = 1.0 if weights is None else float(weights[k])
        v.depth_bias, v.min_view_cosine = depth_bias, min_view_cosine
    return handle, dev, (pos, nrm, uvs, idx), views, (imgs, masks)


def bake_atlas(context, positions, normals, uv, indices, images: Sequence, world_to_clip, camera_positions,
               visibility_masks=None, viewports=None, *, resolution=(1024, 1024), blend_mode="weighted_average",
               visibility_mode="shadow_map", pcf_radius=1, softmax_scale=0.0, mask_floor=0.0,
               weights: Optional[Sequence[float]] = None, depth_bias=1e-3, min_view_cosine=0.05) -> BakeResult:
    import torch

    handle, dev, (pos, nrm, uvs, idx), views, keep = _prepare(
        context, positions, normals, uv, indices, images, world_to_clip, camera_positions, visibility_masks, viewports,
        weights, depth_bias, min_view_cosine)
    n = len(images)
    o = _native.BakeOptions()
    _native.lib().pdr_bake_options_default(ctypes.byref(o))
    o.height, o.width = int(resolution[0]), int(resolution[1])
    o.blend_mode = BLEND_MODES[blend_mode]
    o.visibility_mode = VISIBILITY_MODES[visibility_mode]
    o.pcf_radius, o.softmax_scale, o.mask_floor = pcf_radius, softmax_scale, mask_floor
    H, W = o.height, o.width
    color = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    conf = torch.empty((H, W), dtype=torch.float32, device=dev)
    src = torch.empty((H, W), dtype=torch.int32, device=dev)
    valid = torch.empty((H, W), dtype=torch.float32, device=dev)
    cov = torch.empty((H, W), dtype=torch.float32, device=dev)
    face = torch.empty((H, W), dtype=torch.int32, device=dev)
    used = ctypes.c_int(0)
    _native.check(_native.lib().pdr_bake(handle, pos.data_ptr(), nrm.data_ptr(), uvs.data_ptr(), int(pos.shape[0]),
                                         idx.data_ptr(), int(idx.shape[0]), views, n, ctypes.byref(o),
                                         color.data_ptr(), conf.data_ptr(), src.data_ptr(), valid.data_ptr(),
                                         cov.data_ptr(), face.data_ptr(), ctypes.byref(used)))
    return BakeResult(color, conf, src, valid, cov, face)


def bake_features(context, positions, normals, uv, indices, images: Sequence, world_to_clip, camera_positions,
                  visibility_masks=None, viewports=None, *, resolution=(1024, 1024), visibility_mode="shadow_map",
                  pcf_radius=1, mask_floor=0.0, weights: Optional[Sequence[float]] = None, depth_bias=1e-3,
                  min_view_cosine=0.05):
    """Warm-start features [H, W, 40] (pdr_bake_features): per covered texel the 8
    samples of highest weighted-average confidence, descending, each
    [R, G, B, confidence, cos]; zero elsewhere."""
    import torch

    handle, dev, (pos, nrm, uvs, idx), views, keep = _prepare(
        context, positions, normals, uv, indices, images, world_to_clip, camera_positions, visibility_masks, viewports,
        weights, depth_bias, min_view_cosine)
    o = _native.BakeOptions()
    _native.lib().pdr_bake_options_default(ctypes.byref(o))
    o.height, o.width = int(resolution[0]), int(resolution[1])
    o.blend_mode = 1
    o.visibility_mode = VISIBILITY_MODES[visibility_mode]
    o.pcf_radius, o.mask_floor = pcf_radius, mask_floor
    feats = torch.empty((o.height, o.width, _native.MLP_FEATURES), dtype=torch.float32, device=dev)
    _native.check(_native.lib().pdr_bake_features(handle, pos.data_ptr(), nrm.data_ptr(), uvs.data_ptr(),
                                                  int(pos.shape[0]), idx.data_ptr(), int(idx.shape[0]), views,
                                                  len(images), ctypes.byref(o), feats.data_ptr()))
    return feats


def pad_texture_atlas(color, valid_mask, padding: int, *, coverage_mask=None, fill_unobserved: bool = False,
                      maximum_fill_distance: int = 0) -> np.ndarray:
    """python/photara_drender/baking.py:33-80: a padded copy of ``color`` whose
    valid texels are extended ``padding`` texels into their empty neighbours
    (``valid_mask`` itself is not widened); with ``fill_unobserved`` the covered
    texels (``coverage_mask``) that received no sample take the nearest valid
    colour, up to ``maximum_fill_distance`` texels (0: no limit)."""
    from ._photara_drender import pad_texture_atlas as native_pad

    rgba = np.ascontiguousarray(color, dtype=np.float32)
    valid = np.asarray(valid_mask, dtype=bool)
    # the reference's checks, in its order (each message is part of the interface)
    checks = (
        (lambda: rgba.ndim != 3 or rgba.shape[2] not in (3, 4), "color must have shape [height, width, 3 or 4]"),
        (lambda: valid.shape != rgba.shape[:2], "valid_mask must match the color dimensions"),
        (lambda: padding < 0, "padding must be non-negative"),
    )
    for failed, message in checks:
        if failed():
            raise ValueError(message)
    if not valid.any() or (padding == 0 and not fill_unobserved):
        return rgba.copy()  # nothing to extend from, or nothing asked for
    coverage = None if coverage_mask is None else np.ascontiguousarray(coverage_mask, dtype=np.float32)
    if coverage is not None and coverage.shape != valid.shape:
        raise ValueError("coverage_mask must match the color dimensions")
    if maximum_fill_distance < 0:
        raise ValueError("maximum_fill_distance must be non-negative")
    result = native_pad(rgba, valid.astype(np.float32), int(padding), coverage, bool(fill_unobserved),
                        int(maximum_fill_distance))
    return result[0]  # (colour, filled mask, gutter, unobserved, remaining): the colour


def pad_atlas(context, color, valid_mask, margin: int = 8, *, coverage_mask=None, filled_mask=None,
              fill_unobserved: bool = False, maximum_fill_distance: int = 0, valid_threshold: float = 0.5) -> dict:
    """pdr_pad_texture_atlas on device tensors: pads ``color`` [H, W, 3|4] in place
    (and records fills in ``filled_mask`` when given); returns the padding statistics."""
    handle = context.handle if hasattr(context, "handle") else context.context_handle
    o = _native.PaddingOptions(int(margin), int(bool(fill_unobserved)), int(maximum_fill_distance),
                               float(valid_threshold))
    s = _native.PaddingStats()
    if not (color.is_contiguous() and valid_mask.is_contiguous()):
        raise ValueError("color and valid_mask must be contiguous")
    import torch

    _native.check(_native.lib().pdr_pad_texture_atlas(
        handle, color.data_ptr(), int(color.shape[1]), int(color.shape[0]), int(color.shape[2]),
        valid_mask.data_ptr(), None if coverage_mask is None else coverage_mask.contiguous().data_ptr(),
        None if filled_mask is None else filled_mask.data_ptr(), ctypes.byref(o), ctypes.byref(s),
        torch.cuda.current_stream().cuda_stream))
    return {k: getattr(s, k) for k, _ in _native.PaddingStats._fields_}
