"""ctypes front end of the CPU oracle (``oracle/_build/libpdr_oracle.so``).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference's raster /
interpolate / texture / refine path on the CPU (see ``pdr_oracle.cpp`` for the
reference file:line of every function).  Only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, as the checker; the
product package never does.

Functions take and return numpy arrays with the reference's layouts and raise
the reference's exception classes on error.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PDR_ORACLE_LIB names another build of the same source (the sanitizer build)
_LIB_PATH = os.environ.get("PDR_ORACLE_LIB") or os.path.join(_HERE, "_build", "libpdr_oracle.so")
_lib = None

_F32 = ctypes.POINTER(ctypes.c_float)
_F64 = ctypes.POINTER(ctypes.c_double)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_ERRORS = {1: ValueError, 2: IndexError, 3: OverflowError, 4: RuntimeError}


def build() -> str:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_last_error.restype = ctypes.c_char_p
    return _lib


def _p(a, ptr):
    if a is None:
        return None
    return a.ctypes.data_as(ptr)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _check(rc: int) -> None:
    if rc:
        raise _ERRORS.get(rc, RuntimeError)(lib().orc_last_error().decode())


def set_num_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def build_tile_bins(pos, idx, width, height, viewport=None):
    pos, idx = _f32(pos), _u32(idx)
    tiles = ((width + 15) // 16) * ((height + 15) // 16)
    offsets = np.zeros(tiles + 1, np.uint32)
    vp = None if viewport is None else _f32(viewport)
    _check(lib().orc_build_tile_bins(_p(pos, _F32), pos.shape[0], _p(idx, _U32), idx.shape[0], width, height,
                                     _p(vp, _F32), _p(offsets, _U32), None))
    lst = np.zeros(int(offsets[-1]), np.uint32)
    _check(lib().orc_build_tile_bins(_p(pos, _F32), pos.shape[0], _p(idx, _U32), idx.shape[0], width, height,
                                     _p(vp, _F32), _p(offsets, _U32), _p(lst, _U32)))
    return offsets, lst


def rasterize_forward(pos, idx, resolution, cull_mode=0, want_deriv=True, viewport=None):
    """Returns (raster[H,W,4], deriv[H,W,4] or None, viewport[4]); resolution is (height, width)."""
    pos, idx = _f32(pos), _u32(idx)
    h, w = int(resolution[0]), int(resolution[1])
    raster = np.empty((h, w, 4), np.float32)
    deriv = np.empty((h, w, 4), np.float32) if want_deriv else None
    vp = None if viewport is None else _f32(viewport)
    vp_out = np.zeros(4, np.float32)
    _check(lib().orc_rasterize_forward(_p(pos, _F32), pos.shape[0], _p(idx, _U32), idx.shape[0], w, h,
                                       int(cull_mode), int(bool(want_deriv)), _p(vp, _F32), _p(raster, _F32),
                                       _p(deriv, _F32), _p(vp_out, _F32)))
    return raster, deriv, vp_out


def rasterize_backward(pos, idx, raster, grad_raster, viewport=None):
    pos, idx, raster, grad_raster = _f32(pos), _u32(idx), _f32(raster), _f32(grad_raster)
    h, w = raster.shape[:2]
    out = np.empty((pos.shape[0], 4), np.float32)
    vp = None if viewport is None else _f32(viewport)
    _check(lib().orc_rasterize_backward(_p(pos, _F32), pos.shape[0], _p(idx, _U32), idx.shape[0], _p(raster, _F32),
                                        w, h, _p(grad_raster, _F32), _p(vp, _F32), _p(out, _F32)))
    return out


def interpolate_forward(attr, idx, raster):
    attr, idx, raster = _f32(attr), _u32(idx), _f32(raster)
    h, w = raster.shape[:2]
    a = attr.shape[1]
    out = np.empty((h, w, a), np.float32)
    _check(lib().orc_interpolate_forward(_p(attr, _F32), attr.shape[0], a, _p(idx, _U32), idx.shape[0],
                                         _p(raster, _F32), h * w, _p(out, _F32)))
    return out


def interpolate_backward(attr, idx, raster, grad, reference_mode=False):
    attr, idx, raster, grad = _f32(attr), _u32(idx), _f32(raster), _f32(grad)
    h, w = raster.shape[:2]
    ga = np.empty_like(attr)
    gr = np.empty((h, w, 4), np.float32)
    _check(lib().orc_interpolate_backward(_p(attr, _F32), attr.shape[0], attr.shape[1], _p(idx, _U32), idx.shape[0],
                                          _p(raster, _F32), h * w, _p(grad, _F32), _p(ga, _F32), _p(gr, _F32),
                                          int(bool(reference_mode))))
    return ga, gr


def interpolate_grad_attributes_reference(V, A, idx, raster, grad, outputs=None):
    """The reference's O(V*A*P) attribute-gradient loop (one output per (vertex,
    attribute), every pixel visited) for the flattened outputs in ``outputs``
    (a range, default all V*A); returns them as a flat float32 array."""
    idx, raster, grad = _u32(idx), _f32(raster), _f32(grad)
    o = range(V * A) if outputs is None else outputs
    out = np.empty(max(len(o), 0), np.float32)
    P = raster.shape[0] * raster.shape[1] if raster.ndim == 3 else raster.shape[0]
    lib().orc_interpolate_grad_attributes_range(
        ctypes.c_uint32(V), ctypes.c_uint32(A), _p(idx, _U32), ctypes.c_uint32(idx.shape[0]), _p(raster, _F32),
        ctypes.c_uint32(P), _p(grad, _F32), ctypes.c_uint64(o.start), ctypes.c_uint64(o.stop), _p(out, _F32))
    return out


ADDRESS_MODES = {"clamp": 0, "wrap": 1, "mirror": 2}


def texture_forward(tex, uv, raster, address_mode="clamp"):
    tex, uv, raster = _f32(tex), _f32(uv), _f32(raster)
    th, tw, c = tex.shape
    h, w = raster.shape[:2]
    out = np.empty((h, w, c), np.float32)
    _check(lib().orc_texture_forward(_p(tex, _F32), tw, th, c, ADDRESS_MODES[address_mode], _p(uv, _F32),
                                     _p(raster, _F32), h * w, _p(out, _F32)))
    return out


def texture_backward(tex, uv, raster, grad, address_mode="clamp", compute_uv_gradient=True):
    tex, uv, raster, grad = _f32(tex), _f32(uv), _f32(raster), _f32(grad)
    th, tw, c = tex.shape
    h, w = raster.shape[:2]
    gt = np.empty_like(tex)
    gu = np.empty((h, w, 2), np.float32) if compute_uv_gradient else None
    _check(lib().orc_texture_backward(_p(tex, _F32), tw, th, c, ADDRESS_MODES[address_mode], _p(uv, _F32),
                                      _p(raster, _F32), h * w, _p(grad, _F32), int(bool(compute_uv_gradient)),
                                      _p(gt, _F32), _p(gu, _F32)))
    return gt, gu


def transform_positions(pos3, m16):
    pos3 = _f32(pos3)
    m = _f32(np.asarray(m16).reshape(16))
    out = np.empty((pos3.shape[0], 4), np.float32)
    lib().orc_transform_positions(_p(pos3, _F32), pos3.shape[0], _p(m, _F32), _p(out, _F32))
    return out


def photometric_gradient(tex, uv, target, mask, eps, norm, grad=None, loss=0.0):
    tex, uv, target, mask = _f32(tex), _f32(uv), _f32(target), _f32(mask)
    th, tw = tex.shape[:2]
    grad = np.zeros(tex.shape, np.float32) if grad is None else grad
    lossv = ctypes.c_float(loss)
    lib().orc_photometric_gradient(_p(tex, _F32), tw, th, _p(uv, _F32), _p(target, _F32), _p(mask, _F32),
                                   mask.size, ctypes.c_float(eps), ctypes.c_float(norm), _p(grad, _F32),
                                   ctypes.byref(lossv))
    return grad, lossv.value


def seam_gradient(tex, pairs, eps, norm, grad=None, loss=0.0):
    tex, pairs = _f32(tex), _f32(pairs)
    th, tw = tex.shape[:2]
    grad = np.zeros(tex.shape, np.float32) if grad is None else grad
    lossv = ctypes.c_float(loss)
    lib().orc_seam_gradient(_p(tex, _F32), tw, th, _p(pairs, _F32), pairs.size // 4, ctypes.c_float(eps),
                            ctypes.c_float(norm), _p(grad, _F32), ctypes.byref(lossv))
    return grad, lossv.value


def adam_update(tex, grad, m, v, lr, b1, b2, eps, bc1, bc2, cmin, cmax):
    """In-place on float32 contiguous arrays (tex, grad, m, v)."""
    f = ctypes.c_float
    lib().orc_adam_update(tex.size, f(lr), f(b1), f(b2), f(eps), f(bc1), f(bc2), f(cmin), f(cmax), _p(tex, _F32),
                          _p(grad, _F32), _p(m, _F32), _p(v, _F32))


def refine_schedule(steps, batch, view_valid_weight, lr, min_lr, b1=0.9, b2=0.999, seam_steps=0):
    vw = _f64(view_valid_weight)
    out = {k: np.zeros(steps, np.float32) for k in ("lr", "bc1", "bc2", "norm")}
    out["view_idx"] = np.zeros(steps * batch, np.uint32)
    out["seam_bc1"] = np.zeros(max(seam_steps, 1), np.float32)
    out["seam_bc2"] = np.zeros(max(seam_steps, 1), np.float32)
    f = ctypes.c_float
    lib().orc_refine_schedule(steps, batch, vw.size, _p(vw, _F64), f(lr), f(min_lr), f(b1), f(b2), seam_steps,
                              _p(out["lr"], _F32), _p(out["bc1"], _F32), _p(out["bc2"], _F32), _p(out["norm"], _F32),
                              _p(out["view_idx"], _U32), _p(out["seam_bc1"], _F32), _p(out["seam_bc2"], _F32))
    out["seam_bc1"] = out["seam_bc1"][:seam_steps]
    out["seam_bc2"] = out["seam_bc2"][:seam_steps]
    return out


def refine_view_inputs(image, raster, visibility_mask=None):
    image, raster = _f32(image), _f32(raster)
    p = raster.shape[0] * raster.shape[1]
    target = np.empty((p, 3), np.float32)
    mask = np.empty(p, np.float32)
    vis = None if visibility_mask is None else _f32(visibility_mask)
    wsum = ctypes.c_double(0.0)
    lib().orc_refine_view_inputs(_p(image, _F32), image.shape[2], _p(vis, _F32), _p(raster, _F32), p,
                                 _p(target, _F32), _p(mask, _F32), ctypes.byref(wsum))
    return target, mask, wsum.value


def refine(initial_texture, positions, uv, indices, images, world_to_clip, visibility_masks=None, viewports=None,
           seam_uv_pairs=None, steps=1000, batch_size=4, learning_rate=5e-3, minimum_learning_rate=2.5e-4,
           photometric_epsilon=1e-3, seam_polish_steps=30, seam_learning_rate=1e-3, adam_beta1=0.9,
           adam_beta2=0.999, adam_epsilon=1e-8, clamp_min=0.0, clamp_max=1.0, seam_epsilon=1e-4):
    """src/texture_refiner.cpp:174-454 on the CPU, pixel-sequential gradient order."""
    tex = _f32(initial_texture).copy()
    th, tw = tex.shape[:2]
    pos3 = _f32(positions)
    cached = []
    for k, image in enumerate(images):
        clip = transform_positions(pos3, world_to_clip[k])
        vp = None if viewports is None or viewports[k] is None else viewports[k]
        raster, _, _ = rasterize_forward(clip, indices, image.shape[:2], want_deriv=False, viewport=vp)
        iuv = interpolate_forward(uv, indices, raster)
        vis = None if visibility_masks is None else visibility_masks[k]
        target, mask, wsum = refine_view_inputs(image, raster, vis)
        cached.append((iuv.reshape(-1, 2), target, mask, wsum))
    pairs = None if seam_uv_pairs is None else _f32(seam_uv_pairs).reshape(-1, 4)
    npairs = 0 if pairs is None else pairs.shape[0]
    seam_steps = seam_polish_steps if npairs else 0
    sched = refine_schedule(steps, batch_size, [c[3] for c in cached], learning_rate, minimum_learning_rate,
                            adam_beta1, adam_beta2, seam_steps)
    grad = np.zeros_like(tex)
    m = np.zeros_like(tex)
    v = np.zeros_like(tex)
    losses = np.zeros(steps, np.float32)
    for s in range(steps):
        loss = 0.0
        for b in range(batch_size):
            iuv, target, mask, _ = cached[int(sched["view_idx"][s * batch_size + b])]
            _, loss = photometric_gradient(tex, iuv, target, mask, photometric_epsilon, float(sched["norm"][s]),
                                           grad, loss)
        losses[s] = loss
        adam_update(tex, grad, m, v, float(sched["lr"][s]), adam_beta1, adam_beta2, adam_epsilon,
                    float(sched["bc1"][s]), float(sched["bc2"][s]), clamp_min, clamp_max)
    seam_losses = np.zeros(seam_steps, np.float32)
    if seam_steps:
        grad[:] = 0
        m[:] = 0
        v[:] = 0
        norm = float(np.float32(1.0) / np.float32(npairs * 3))
        for k in range(seam_steps):
            _, seam_losses[k] = seam_gradient(tex, pairs, seam_epsilon, norm, grad, 0.0)
            adam_update(tex, grad, m, v, seam_learning_rate, adam_beta1, adam_beta2, adam_epsilon,
                        float(sched["seam_bc1"][k]), float(sched["seam_bc2"][k]), clamp_min, clamp_max)
    return tex, losses, seam_losses


def raster_fixed_f64(pos, idx, raster_ids, viewport=None):
    pos, idx, raster_ids = _f64(pos), _u32(idx), _f32(raster_ids)
    h, w = raster_ids.shape[:2]
    out = np.empty((h, w, 4), np.float64)
    vp = None if viewport is None else _f64(viewport)
    lib().orc_raster_fixed_f64(_p(pos, _F64), pos.shape[0], _p(idx, _U32), idx.shape[0], _p(raster_ids, _F32), w, h,
                               _p(vp, _F64), _p(out, _F64))
    return out


def interpolate_forward_f64(attr, idx, raster):
    attr, idx, raster = _f64(attr), _u32(idx), _f64(raster)
    h, w = raster.shape[:2]
    out = np.empty((h, w, attr.shape[1]), np.float64)
    lib().orc_interpolate_forward_f64(_p(attr, _F64), attr.shape[0], attr.shape[1], _p(idx, _U32), idx.shape[0],
                                      _p(raster, _F64), h * w, _p(out, _F64))
    return out


def texture_forward_f64(tex, uv, raster, address_mode="clamp"):
    tex, uv, raster = _f64(tex), _f64(uv), _f64(raster)
    th, tw, c = tex.shape
    h, w = raster.shape[:2]
    out = np.empty((h, w, c), np.float64)
    lib().orc_texture_forward_f64(_p(tex, _F64), tw, th, c, ADDRESS_MODES[address_mode], _p(uv, _F64),
                                  _p(raster, _F64), h * w, _p(out, _F64))
    return out


class _BakeView(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("channels", ctypes.c_uint32),
                ("image", ctypes.c_void_p), ("mask", ctypes.c_void_p), ("world_to_clip", ctypes.c_float * 16),
                ("camera", ctypes.c_float * 3), ("viewport", ctypes.c_float * 4), ("weight", ctypes.c_float),
                ("depth_bias", ctypes.c_float), ("min_view_cosine", ctypes.c_float)]


def suggest_softmax_scale(positions, camera_positions):
    """src/texture_padding.cpp:187-222"""
    pos, cams = _f32(positions), _f32(camera_positions)
    fn = lib().orc_suggest_softmax_scale
    fn.restype = ctypes.c_float
    return float(fn(_p(pos, _F32), pos.shape[0], _p(cams, _F32), cams.shape[0]))


BLEND_MODES = {"best_view": 0, "weighted_average": 1, "softmax": 2}


def _bake_views(positions, images, world_to_clip, camera_positions, visibility_masks, viewports, mode,
                softmax_scale, mask_floor, weights, depth_bias, min_view_cosine):
    """orc_bake_view records of the views (and the arrays they point into)."""
    pos = _f32(positions)
    n = len(images)
    mats = _f32(np.asarray(world_to_clip).reshape(n, 16))
    cams = _f32(np.asarray(camera_positions).reshape(n, 3))
    scale = float(softmax_scale)
    if mode == 2 and not (scale > 0 and np.isfinite(scale)):
        scale = suggest_softmax_scale(pos, cams)
        if not (scale > 0 and np.isfinite(scale)):
            scale = 1.0
    keep, views = [], (_BakeView * n)()
    for k in range(n):
        img = _f32(images[k])
        h, w, c = img.shape
        v = views[k]
        v.width, v.height, v.channels = w, h, c
        v.image = img.ctypes.data
        keep.append(img)
        m = None if visibility_masks is None else visibility_masks[k]
        if m is not None:
            m = _f32(m)
            if mask_floor > 0:
                m = np.ascontiguousarray(np.maximum(m, np.float32(mask_floor)), np.float32)
            keep.append(m)
            v.mask = m.ctypes.data
        v.world_to_clip = (ctypes.c_float * 16)(*mats[k].tolist())
        v.camera = (ctypes.c_float * 3)(*cams[k].tolist())
        vp = None if viewports is None else viewports[k]
        v.viewport = (ctypes.c_float * 4)(*([0.0, 0.0, float(w), float(h)] if vp is None else [float(x) for x in vp]))
        wk = 1.0 if weights is None else float(weights[k])
        v.weight = wk * scale if mode == 2 else wk
        v.depth_bias, v.min_view_cosine = depth_bias, min_view_cosine
    return views, keep


def bake(positions, normals, uv, indices, images, world_to_clip, camera_positions, visibility_masks=None,
         viewports=None, resolution=(1024, 1024), blend_mode="weighted_average", pcf_radius=1, softmax_scale=0.0,
         mask_floor=0.0, weights=None, depth_bias=1e-3, min_view_cosine=0.05):
    """TextureBaker::bake (src/texture_baker.cpp:156-361) with shadow-map visibility.
    Returns (color [H,W,4], confidence, source_view (u32), valid, coverage, face_ids (u32))."""
    pos, nrm, uvs, idx = _f32(positions), _f32(normals), _f32(uv), _u32(indices)
    n = len(images)
    mode = BLEND_MODES[blend_mode]
    views, keep = _bake_views(pos, images, world_to_clip, camera_positions, visibility_masks, viewports, mode,
                              softmax_scale, mask_floor, weights, depth_bias, min_view_cosine)
    H, W = resolution
    color = np.empty((H, W, 4), np.float32)
    conf = np.empty((H, W), np.float32)
    src = np.empty((H, W), np.uint32)
    valid = np.empty((H, W), np.float32)
    cov = np.empty((H, W), np.float32)
    face = np.empty((H, W), np.uint32)
    _check(lib().orc_bake(_p(pos, _F32), _p(nrm, _F32), _p(uvs, _F32), pos.shape[0], _p(idx, _U32), idx.shape[0],
                          views, n, W, H, mode, pcf_radius, _p(color, _F32), _p(conf, _F32), _p(src, _U32),
                          _p(valid, _F32), _p(cov, _F32), _p(face, _U32)))
    return color, conf, src, valid, cov, face


def bake_features(positions, normals, uv, indices, images, world_to_clip, camera_positions, visibility_masks=None,
                  viewports=None, resolution=(1024, 1024), pcf_radius=1, mask_floor=0.0, weights=None,
                  depth_bias=1e-3, min_view_cosine=0.05):
    """Warm-start features [H, W, 40]: per covered texel the 8 samples of highest
    weighted-average confidence (project_texture.hlsl:208), descending, ties in
    view order, each [R, G, B, confidence, cos]; zero elsewhere."""
    pos, nrm, uvs, idx = _f32(positions), _f32(normals), _f32(uv), _u32(indices)
    n = len(images)
    views, keep = _bake_views(pos, images, world_to_clip, camera_positions, visibility_masks, viewports, 1, 0.0,
                              mask_floor, weights, depth_bias, min_view_cosine)
    H, W = resolution
    feats = np.empty((H, W, 40), np.float32)
    _check(lib().orc_bake_features(_p(pos, _F32), _p(nrm, _F32), _p(uvs, _F32), pos.shape[0], _p(idx, _U32),
                                   idx.shape[0], views, n, W, H, pcf_radius, _p(feats, _F32)))
    return feats


def pad_texture_atlas(color, valid_mask, margin=4, coverage_mask=None, filled_mask=None, fill_unobserved=True,
                      maximum_fill_distance=0, valid_threshold=0.5):
    """pad_texture_atlas (src/texture_padding.cpp:226-351, TexturePaddingOptions
    defaults of include/photara_drender/texture_padding.hpp:33-44).  Returns
    (color, filled_mask, stats) with stats keys valid_texels, gutter_texels,
    unobserved_texels, remaining_unobserved, filled_texels."""
    color = np.array(color, np.float32, copy=True, order="C")
    H, W, C = color.shape
    valid = _f32(valid_mask)
    cov = None if coverage_mask is None else _f32(coverage_mask)
    filled = np.zeros((H, W), np.float32) if filled_mask is None else np.array(filled_mask, np.float32, copy=True)
    st = np.zeros(5, np.uint64)
    _check(lib().orc_pad_texture_atlas(_p(color, _F32), W, H, C, _p(valid, _F32),
                                       None if cov is None else _p(cov, _F32), _p(filled, _F32), int(margin),
                                       int(bool(fill_unobserved)), int(maximum_fill_distance),
                                       ctypes.c_float(valid_threshold), st.ctypes.data_as(ctypes.c_void_p)))
    keys = ("valid_texels", "gutter_texels", "unobserved_texels", "remaining_unobserved", "filled_texels")
    return color, filled, {k: int(v) for k, v in zip(keys, st)}
