"""Drives libpdr's C ABI from host (numpy) arrays, for the sanitizer runs.

    python tests/hostsan_driver.py cpu|gpu|planted

Loaded with the library named by PDR_LIB (the host-sanitizer build) and
clang's AddressSanitizer runtime preloaded: a sanitizer report aborts the
process with a non-zero status.  ``cpu`` runs what needs no GPU (the
library loads, options defaults, host-only helpers, argument errors,
context creation failing without a device); ``gpu`` also runs small
workloads through every host-array path (staging, batched views, autotune
cache text, timing, refine with seams / TV / checkpoints, bake, padding)
and checks the raster against the oracle.  Prints one line per step and
"hostsan driver: ok" at the end.  ``planted`` reads past a host array on
purpose (CPU only), to show that the sanitizer reports it.
"""
import ctypes
import importlib.util
import os
import sys
import types

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.modules["torch"] = None  # host arrays only: the library runs on HIP directly

_pkg = types.ModuleType("pdr_hostsan")
_pkg.__path__ = []
sys.modules["pdr_hostsan"] = _pkg


def _load(name):
    spec = importlib.util.spec_from_file_location(f"pdr_hostsan.{name}",
                                                  os.path.join(ROOT, "photara_drender_amd", f"{name}.py"))
    m = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = m
    spec.loader.exec_module(m)
    setattr(_pkg, name, m)
    return m


native = _load("_native")
scenes = _load("scenes")
refine = _load("refine")
F = ctypes.c_float


def step(msg):
    print(msg, flush=True)


def p(a):
    return None if a is None else a.ctypes.data


def expect_error(status, kind, text):
    assert status != 0, f"expected an error containing {text!r}"
    msg = native.lib().pdr_last_error().decode()
    err = native._ERRORS.get(status, RuntimeError)
    assert err is kind and text in msg, (status, msg)


class HipMemory:
    """Device arrays through the HIP runtime's C API (no torch in this process)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def alloc(self, nbytes):
        ptr = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) == 0
        self.ptrs.append(ptr)
        return ptr.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        ptr = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(ctypes.c_void_p(ptr), ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes),
                                 1) == 0  # host to device
        return ptr

    def download(self, ptr, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert self.rt.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes),
                                 2) == 0  # device to host
        return out

    def free_all(self):
        for ptr in self.ptrs:
            self.rt.hipFree(ptr)
        self.ptrs = []


def cpu_part():
    lib = native.lib()
    step(f"loaded {native.LIB_PATH}")
    o = native.RefineOptions()
    lib.pdr_refine_options_default(ctypes.byref(o))
    assert o.steps == 1000 and o.batch_size == 4
    b = native.BakeOptions()
    lib.pdr_bake_options_default(ctypes.byref(b))
    v = native.BakeView()
    lib.pdr_bake_view_default(ctypes.byref(v))
    po = native.PaddingOptions()
    lib.pdr_padding_options_default(ctypes.byref(po))
    pos = np.array([[0, 0, 1], [0, 0, 3]], np.float32)
    cams = np.array([[0, 0, 0]], np.float32)
    s = lib.pdr_suggest_softmax_scale(p(pos), 2, p(cams), 1)
    assert abs(s - 20.0) < 1e-4, s
    step("defaults and host helpers")
    tri = np.array([[-0.5, -0.5, 0, 1], [0.5, -0.5, 0, 1], [0, 0.5, 0, 1]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    out = np.zeros((4, 4, 4), np.float32)
    expect_error(lib.pdr_rasterize_forward(None, p(tri), 3, p(idx), 1, 4, 4, 0, 0, None, p(out), None, None, None),
                 ValueError, "context is null")
    expect_error(lib.pdr_autotune_import(None, b"forward 1 2\n"), ValueError, "context is null")
    h = ctypes.c_void_p()
    rc = lib.pdr_context_create(0, 0, ctypes.byref(h))
    if rc == 0:
        lib.pdr_context_destroy(h)
        step("context created (a device is present)")
    else:
        assert lib.pdr_last_error().decode()
        step("context creation without a device reports an error")


def gpu_part():
    import oracle

    lib = native.lib()
    ctx = native.Context(0, flags=native.PDR_FLAG_VALIDATE_DEVICE_INPUTS)
    sc = scenes.chain_scene(level=3, res=(48, 64), tex_size=32, channels=4, seed=3)
    H, W = sc.resolution
    V, T = sc.clip.shape[0], sc.indices.shape[0]
    idx = np.ascontiguousarray(sc.indices, np.uint32)
    raster = np.empty((H, W, 4), np.float32)
    deriv = np.empty((H, W, 4), np.float32)
    vp = (F * 4)()
    native.check(lib.pdr_rasterize_forward(ctx.handle, p(sc.clip), V, p(idx), T, W, H, 0, 1, None, p(raster),
                                           p(deriv), vp, None))
    ref, ref_d, _ = oracle.rasterize_forward(sc.clip, sc.indices, (H, W))
    assert np.array_equal(raster.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(deriv.view(np.uint32), ref_d.view(np.uint32))
    step("rasterize forward (host arrays) = oracle")
    gr = np.random.default_rng(1).standard_normal((H, W, 4)).astype(np.float32)
    gpos = np.empty((V, 4), np.float32)
    native.check(lib.pdr_rasterize_backward(ctx.handle, p(sc.clip), V, p(idx), T, p(raster), W, H, p(gr), None,
                                            p(gpos), None))
    uvs = np.empty((H, W, 2), np.float32)
    native.check(lib.pdr_interpolate_forward(ctx.handle, p(sc.uv), V, 2, p(idx), T, p(raster), W, H, p(uvs), None))
    guv = np.empty((V, 2), np.float32)
    grr = np.empty((H, W, 4), np.float32)
    native.check(lib.pdr_interpolate_backward(ctx.handle, p(sc.uv), V, 2, p(idx), T, p(raster), W, H, p(uvs),
                                              p(guv), p(grr), None))
    th, tw, C = sc.texture.shape
    color = np.empty((H, W, C), np.float32)
    gtex = np.empty_like(sc.texture)
    guvs = np.empty((H, W, 2), np.float32)
    for mode in range(3):
        native.check(lib.pdr_texture_forward(ctx.handle, p(sc.texture), tw, th, C, mode, p(uvs), p(raster), W, H,
                                             p(color), None))
        native.check(lib.pdr_texture_backward(ctx.handle, p(sc.texture), tw, th, C, mode, p(uvs), p(raster), W, H,
                                              p(color), 1, p(gtex), p(guvs), None))
    step("raster backward, interpolate, texture (three address modes)")
    # batched views of different sizes and the no-allocation forward (device memory only)
    hip = HipMemory()
    d_clip, d_idx = hip.upload(sc.clip), hip.upload(idx)
    sizes = [(32, 40), (20, 24)]
    rv = (native.RasterView * 2)()
    outs = []
    for k, (h, w) in enumerate(sizes):
        g = np.random.default_rng(k).standard_normal((h, w, 4)).astype(np.float32)
        r, d, dg, gp = hip.alloc(h * w * 16), hip.alloc(h * w * 16), hip.upload(g), hip.alloc(V * 16)
        outs.append((h, w, r, gp))
        rv[k].positions, rv[k].width, rv[k].height = d_clip, w, h
        rv[k].viewport = (F * 4)(0, 0, 0, 0)
        rv[k].raster, rv[k].derivatives, rv[k].grad_raster, rv[k].grad_positions = r, d, dg, gp
    native.check(lib.pdr_rasterize_forward_batched(ctx.handle, rv, 2, V, d_idx, T, 0, 1, None))
    native.check(lib.pdr_rasterize_backward_batched(ctx.handle, rv, 2, V, d_idx, T, None))
    native.check(lib.pdr_synchronize(ctx.handle, None))
    for h, w, r, _ in outs:
        ref, _, _ = oracle.rasterize_forward(sc.clip, sc.indices, (h, w))
        got = hip.download(r, (h, w, 4))
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    na = hip.alloc(H * W * 16)
    native.check(lib.pdr_rasterize_forward_noalloc(ctx.handle, d_clip, V, d_idx, T, W, H, 0, 0, None, na, None, None,
                                                   None))
    native.check(lib.pdr_synchronize(ctx.handle, None))
    assert np.array_equal(hip.download(na, (H, W, 4)).view(np.uint32), raster.view(np.uint32))
    hip.free_all()
    step("batched raster forward + backward and the no-allocation forward (device arrays)")
    # argument errors with a context
    bad = idx.copy()
    bad[0, 0] = V + 5
    with_raise = native.Context(0, flags=native.PDR_FLAG_VALIDATE_DEVICE_INPUTS)
    expect_error(lib.pdr_rasterize_forward(with_raise.handle, p(sc.clip), V, p(bad), T, W, H, 0, 1, None, p(raster),
                                           p(deriv), vp, None), IndexError, "")
    expect_error(lib.pdr_rasterize_forward(ctx.handle, p(sc.clip), V, p(idx), T, 0, H, 0, 1, None, p(raster), None,
                                           None, None), ValueError, "")
    expect_error(lib.pdr_texture_forward(ctx.handle, p(sc.texture), 0, th, C, 0, p(uvs), p(raster), W, H, p(color),
                                         None), ValueError, "")
    step("argument errors")
    # autotune cache text, timing
    tuned = native.Context(0, flags=native.PDR_FLAG_AUTOTUNE)
    native.check(lib.pdr_rasterize_forward(tuned.handle, p(sc.clip), V, p(idx), T, W, H, 0, 1, None, p(raster),
                                           p(deriv), vp, None))
    text = tuned.autotune_export()
    assert text.startswith("forward_out"), text
    tuned.autotune_clear()
    tuned.autotune_import(text + "\n")
    assert tuned.autotune_export() == text
    for line in ("forward_out 1 2\n", "sideways 1 2 3 4 5 6 7 8 9\n", "forward_out 1 2 3 x 5 6 7 8 9\n",
                 "forward_out 1 2 3 4096 4096 512 4096 7 4\n"):
        try:
            tuned.autotune_import(line)
            raise AssertionError(f"accepted {line!r}")
        except ValueError:
            pass
    tuned.set_timing(True)
    native.check(lib.pdr_rasterize_forward(tuned.handle, p(sc.clip), V, p(idx), T, W, H, 0, 1, None, p(raster),
                                           p(deriv), vp, None))
    ms, n = tuned.kernel_time("raster_candidates")
    assert n >= 1
    tuned.set_timing(False)
    step("autotune cache text and timing")
    # refine: a quad atlas seen by two views, seams, TV over charts, checkpoints
    q = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    quv = np.array([[0.05, 0.05], [0.95, 0.05], [0.95, 0.95], [0.05, 0.95]], np.float32)
    qidx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    mats = np.stack([scenes.world_to_clip((0.3, 0.2, 3.0), 1.25), scenes.world_to_clip((-0.4, 0.1, 2.8), 1.25)])
    rng = np.random.default_rng(5)
    imgs = [rng.uniform(0, 1, (40, 50, 3)).astype(np.float32) for _ in range(2)]
    masks = [(rng.uniform(0, 1, (40, 50)) > 0.2).astype(np.float32) for _ in range(2)]
    init = np.full((32, 32, 3), 0.5, np.float32)
    charts = np.zeros((32, 32), np.int32)
    charts[:, 16:] = 1
    pairs = rng.uniform(0.1, 0.9, (20, 2, 2)).astype(np.float32)
    res = refine.refine_atlas(ctx, init, q, quv, qidx, imgs, mats, masks, seam_uv_pairs=pairs, steps=12,
                              batch_size=2, seam_polish_steps=3, tv_weight=1e-3, chart_ids=charts,
                              joint_seam_weight=1e-2, checkpoint_every=4)
    assert np.isfinite(res.color).all() and res.loss_history.shape == (12,) and res.checkpoints.shape[0] == 3
    assert res.seam_loss_history.shape == (3,) and np.isfinite(res.seam_loss_history).all()
    res2 = refine.refine_atlas(ctx, init, q, quv, qidx, imgs, mats, steps=6, batch_size=4, use_graph=False)
    assert np.isfinite(res2.color).all()
    step("refine (graph and eager, seams, TV, checkpoints)")
    # bake: one view of the quad, host arrays in and out
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    views = (native.BakeView * 2)()
    for k in range(2):
        v = views[k]
        lib.pdr_bake_view_default(ctypes.byref(v))
        v.height, v.width, v.channels = 40, 50, 3
        v.image, v.visibility_mask = p(imgs[k]), p(masks[k])
        v.world_to_clip = (F * 16)(*mats[k].ravel().tolist())
        v.camera_position = (F * 3)(*([0.3, 0.2, 3.0] if k == 0 else [-0.4, 0.1, 2.8]))
    bo = native.BakeOptions()
    lib.pdr_bake_options_default(ctypes.byref(bo))
    bo.width = bo.height = 32
    outs = [np.empty((32, 32, 4), np.float32), np.empty((32, 32), np.float32), np.empty((32, 32), np.uint32),
            np.empty((32, 32), np.float32), np.empty((32, 32), np.float32), np.empty((32, 32), np.uint32)]
    used = ctypes.c_int(0)
    native.check(lib.pdr_bake(ctx.handle, p(q), p(nrm), p(quv), 4, p(qidx), 2, views, 2, ctypes.byref(bo),
                              *[p(a) for a in outs], ctypes.byref(used)))
    feats = np.empty((32, 32, native.MLP_FEATURES), np.float32)
    native.check(lib.pdr_bake_features(ctx.handle, p(q), p(nrm), p(quv), 4, p(qidx), 2, views, 2, ctypes.byref(bo),
                                       p(feats)))
    step("bake and bake features")
    # padding of the baked atlas
    col = outs[0].copy()
    filled = np.zeros((32, 32), np.float32)
    po = native.PaddingOptions(4, 1, 0, 0.5)
    st = native.PaddingStats()
    native.check(lib.pdr_pad_texture_atlas(ctx.handle, p(col), 32, 32, 4, p(outs[3]), p(outs[4]), p(filled),
                                           ctypes.byref(po), ctypes.byref(st), None))
    step("padding")
    for c in (ctx, with_raise, tuned):
        c.close()


def planted():
    """A host array one row shorter than the count passed (sanitizer self-check:
    the run must end in a heap-buffer-overflow report)."""
    pos = np.ones((200, 3), np.float32)
    cams = np.zeros((1, 3), np.float32)
    native.lib().pdr_suggest_softmax_scale(p(pos), 201, p(cams), 1)


if __name__ == "__main__":
    mode = sys.argv[1] if len(sys.argv) > 1 else "cpu"
    if mode == "planted":
        planted()
        sys.exit(0)
    cpu_part()
    if mode == "gpu":
        gpu_part()
    print("hostsan driver: ok", flush=True)
