"""Atlas refinement with the extended options of ``pdr_refine``
(TextureRefiner::refine, src/texture_refiner.cpp:174-454, plus chart total
variation, a joint seam term, hipGraph capture control and atlas checkpoints).

``TextureBaker.refine_texture`` of the native module is the reference's exact
entry point; this wrapper exposes the same loop for torch tensors already on the
GPU (no host staging) and for the project's extensions.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _native


@dataclass
class RefineResult:
    color: object  # [H, W, 3] (torch tensor on the device of the inputs, or numpy)
    loss_history: np.ndarray
    seam_loss_history: np.ndarray
    precompute_seconds: float
    optimization_seconds: float
    checkpoints: Optional[object] = None


def _ptr(x):
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data


def _contig(x, dtype):
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        import torch

        return x.detach().to(dtype={np.float32: torch.float32, np.int32: torch.int32}[dtype]).contiguous()
    return np.ascontiguousarray(x, dtype=dtype)


def refine_atlas(context, initial_texture, positions, uv, indices, images: Sequence, world_to_clip,
                 visibility_masks=None, viewports=None, seam_uv_pairs=None, *, steps=1000, batch_size=4,
                 learning_rate=5e-3, minimum_learning_rate=2.5e-4, photometric_epsilon=1e-3, seam_polish_steps=30,
                 seam_learning_rate=1e-3, seam_epsilon=1e-4, adam_beta1=0.9, adam_beta2=0.999, adam_epsilon=1e-8,
                 clamp_min=0.0, clamp_max=1.0, tv_weight=0.0, tv_epsilon=1e-4, chart_ids=None,
                 joint_seam_weight=0.0, use_graph=True, checkpoint_every=0, stop_after=0) -> RefineResult:
    """Runs pdr_refine.  Arrays may be numpy (host) or CUDA tensors (device).

    ``context`` is a ``_native.Context`` or an object with ``context_handle``.
    ``chart_ids`` [H, W] int32 (< 0 = texel outside every chart) feeds the chart TV term.
    ``stop_after`` > 0 runs only the first ``stop_after`` steps of the ``steps``-step
    schedule (then the seam polish).
    """
    handle = context.handle if hasattr(context, "handle") else context.context_handle
    tex = _contig(initial_texture, np.float32)
    height, width = int(tex.shape[0]), int(tex.shape[1])
    pos = _contig(positions, np.float32)
    uvv = _contig(uv, np.float32)
    idx = indices
    if hasattr(idx, "data_ptr"):
        import torch

        idx = idx.detach().to(torch.int32).contiguous()
    else:
        idx = np.ascontiguousarray(idx, dtype=np.uint32)
    n = len(images)
    imgs = [_contig(im, np.float32) for im in images]
    masks = [None] * n if visibility_masks is None else [_contig(m, np.float32) for m in visibility_masks]
    mats = np.ascontiguousarray(np.asarray(world_to_clip, dtype=np.float32).reshape(n, 16))
    views = (_native.RefineView * n)()
    for k in range(n):
        v = views[k]
        v.height, v.width, v.channels = int(imgs[k].shape[0]), int(imgs[k].shape[1]), int(imgs[k].shape[2])
        v.image = _ptr(imgs[k])
        v.visibility_mask = _ptr(masks[k])
        v.world_to_clip = (ctypes.c_float * 16)(*mats[k].tolist())
        vp = None if viewports is None else viewports[k]
        v.has_viewport = 0 if vp is None else 1
        if vp is not None:
            v.viewport = (ctypes.c_float * 4)(*[float(x) for x in vp])
    o = _native.RefineOptions()
    _native.lib().pdr_refine_options_default(ctypes.byref(o))
    o.width, o.height, o.steps, o.batch_size = width, height, steps, batch_size
    o.learning_rate, o.minimum_learning_rate = learning_rate, minimum_learning_rate
    o.photometric_epsilon, o.adam_beta1, o.adam_beta2, o.adam_epsilon = photometric_epsilon, adam_beta1, adam_beta2, \
        adam_epsilon
    o.clamp_min, o.clamp_max = clamp_min, clamp_max
    o.seam_polish_steps, o.seam_learning_rate, o.seam_epsilon = seam_polish_steps, seam_learning_rate, seam_epsilon
    o.tv_weight, o.tv_epsilon = tv_weight, tv_epsilon
    charts = None if chart_ids is None else _contig(chart_ids, np.int32)
    o.chart_ids = _ptr(charts)
    o.joint_seam_weight = joint_seam_weight
    o.use_graph = int(bool(use_graph))
    o.checkpoint_every = int(checkpoint_every)
    o.stop_after = int(stop_after)
    device = hasattr(tex, "data_ptr")
    ckpt = None
    if checkpoint_every:
        count = steps // checkpoint_every
        if device:
            import torch

            ckpt = torch.empty((count, height, width, 3), dtype=torch.float32, device=tex.device)
        else:
            ckpt = np.empty((count, height, width, 3), np.float32)
    o.checkpoints = _ptr(ckpt)
    pairs = None if seam_uv_pairs is None else _contig(seam_uv_pairs, np.float32)
    npairs = 0 if pairs is None else int(np.prod(pairs.shape)) // 4
    if device:
        import torch

        color = torch.empty((height, width, 3), dtype=torch.float32, device=tex.device)
    else:
        color = np.empty((height, width, 3), np.float32)
    seam_steps = seam_polish_steps if npairs else 0
    loss = np.zeros(steps, np.float32)
    seam_loss = np.zeros(seam_steps, np.float32)
    times = (ctypes.c_double * 2)()
    _native.check(_native.lib().pdr_refine(
        handle, _ptr(tex), _ptr(pos), int(pos.shape[0]), _ptr(uvv), _ptr(idx), int(idx.shape[0]), views, n,
        _ptr(pairs), npairs, ctypes.byref(o), _ptr(color), loss.ctypes.data, seam_loss.ctypes.data if seam_steps else None,
        times))
    return RefineResult(color, loss, seam_loss, times[0], times[1], ckpt)
