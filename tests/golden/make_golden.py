"""Writes the small golden fixtures of tests/golden/ from the CPU oracle.

The reference ships no vectors for this path (its tests assert closed forms,
finite differences and an nvdiffrast comparison that needs CUDA), so these
fixtures are the oracle's outputs on the reference tests' own inputs
(tests/native_tests.cpp:127-233, tests/python_smoke.py:20-32,121-134,
tests/compare_nvdiffrast.py:22-27).  They pin the oracle against regressions
and give the GPU tests a fixed target.  Regenerate with
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

TRIANGLE = np.array([[-0.75, -0.75, 0.0, 1.0], [0.75, -0.75, 0.0, 1.0], [0.0, 0.75, 0.0, 1.0]], np.float32)
ONE = np.array([[0, 1, 2]], np.uint32)


def native_case():
    raster, deriv, _ = oracle.rasterize_forward(TRIANGLE, ONE, (32, 32))
    grad = np.zeros_like(raster)
    grad[..., 0] = (raster[..., 3] > 0).astype(np.float32)
    grad_pos = oracle.rasterize_backward(TRIANGLE, ONE, raster, grad)
    attrs = np.eye(3, dtype=np.float32)
    interp = oracle.interpolate_forward(attrs, ONE, raster)
    ga, gr = oracle.interpolate_backward(attrs, ONE, raster, np.ones_like(interp))
    vuv = np.array([[0.2, 0.2], [0.8, 0.2], [0.5, 0.8]], np.float32)
    uv = oracle.interpolate_forward(vuv, ONE, raster)
    tex = np.zeros((4, 4, 2), np.float32)
    for y in range(4):
        for x in range(4):
            tex[y, x] = (x + y * 2, x * 3 - y)
    sampled = oracle.texture_forward(tex, uv, raster)
    gs = np.zeros_like(sampled)
    first = np.argwhere(raster.reshape(-1, 4)[:, 3] != 0)[0, 0]
    gs.reshape(-1, 2)[first, 0] = 1.0
    gt, gu = oracle.texture_backward(tex, uv, raster, gs)
    vr, vd, _ = oracle.rasterize_forward(TRIANGLE, ONE, (32, 32), viewport=(8.0, 4.0, 16.0, 20.0))
    return dict(positions=TRIANGLE, indices=ONE, raster=raster, deriv=deriv, grad_raster=grad, grad_pos=grad_pos,
                attrs=attrs, interp=interp, grad_attr=ga, grad_raster_interp=gr, vertex_uv=vuv, uv=uv, texture=tex,
                sampled=sampled, grad_sampled=gs, grad_texture=gt, grad_uv=gu, viewport_raster=vr,
                viewport_deriv=vd)


def smoke_case():
    raster, deriv, _ = oracle.rasterize_forward(TRIANGLE, ONE, (16, 20))
    vuv = np.array([[0.2, 0.2], [0.8, 0.2], [0.5, 0.8]], np.float32)
    uv = oracle.interpolate_forward(vuv, ONE, raster)
    tex = (np.arange(48, dtype=np.float32).reshape(4, 4, 3) / 48.0).astype(np.float32)
    out = {"raster": raster, "deriv": deriv, "uv": uv, "texture": tex}
    for mode in ("clamp", "wrap", "mirror"):
        shifted = (uv + 1.25).astype(np.float32)
        out["sampled_" + mode] = oracle.texture_forward(tex, shifted, raster, mode)
        gt, gu = oracle.texture_backward(tex, shifted, raster, np.ones((16, 20, 3), np.float32), mode)
        out["grad_texture_" + mode] = gt
        out["grad_uv_" + mode] = gu
    vr, _, _ = oracle.rasterize_forward(TRIANGLE, ONE, (20, 24), viewport=(4.0, 3.0, 16.0, 12.0))
    out["viewport_raster"] = vr
    return out


def nvdiffrast_case():
    pos = np.array([[-0.75, -0.75, 0.1, 1.0], [0.75, -0.75, 0.2, 1.0], [0.0, 0.75, 0.3, 1.0]], np.float32)
    raster, deriv, _ = oracle.rasterize_forward(pos, ONE, (64, 64))
    g = np.random.default_rng(7).standard_normal(raster.shape, dtype=np.float32)
    return dict(positions=pos, raster=raster, deriv=deriv, grad_raster=g,
                grad_pos=oracle.rasterize_backward(pos, ONE, raster, g))


def main():
    for name, fn in (("native_32", native_case), ("smoke_16x20", smoke_case), ("nvdiffrast_64", nvdiffrast_case)):
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **fn())
        print("wrote", name)


if __name__ == "__main__":
    main()
