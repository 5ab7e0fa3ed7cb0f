"""Golden fixtures (tests/golden/*.npz, written by make_golden.py on the reference
tests' own inputs): the oracle must reproduce them bit for bit (CPU), and the GPU
kernels must match them to the parity bar (GPU)."""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_oracle_reproduces_native_fixture():
    g = load("native_32")
    r, d, _ = oracle.rasterize_forward(g["positions"], g["indices"], (32, 32))
    same_bits(r, g["raster"])
    same_bits(d, g["deriv"])
    same_bits(oracle.rasterize_backward(g["positions"], g["indices"], r, g["grad_raster"]), g["grad_pos"])
    same_bits(oracle.interpolate_forward(g["attrs"], g["indices"], r), g["interp"])
    ga, gr = oracle.interpolate_backward(g["attrs"], g["indices"], r, np.ones_like(g["interp"]))
    same_bits(ga, g["grad_attr"])
    same_bits(gr, g["grad_raster_interp"])
    same_bits(oracle.texture_forward(g["texture"], g["uv"], r), g["sampled"])
    gt, gu = oracle.texture_backward(g["texture"], g["uv"], r, g["grad_sampled"])
    same_bits(gt, g["grad_texture"])
    same_bits(gu, g["grad_uv"])
    vr, vd, _ = oracle.rasterize_forward(g["positions"], g["indices"], (32, 32), viewport=(8.0, 4.0, 16.0, 20.0))
    same_bits(vr, g["viewport_raster"])


def test_oracle_reproduces_smoke_and_nvdiffrast_fixtures():
    g = load("smoke_16x20")
    tri = load("native_32")
    r, d, _ = oracle.rasterize_forward(tri["positions"], tri["indices"], (16, 20))
    same_bits(r, g["raster"])
    for mode in ("clamp", "wrap", "mirror"):
        shifted = (g["uv"] + 1.25).astype(np.float32)
        same_bits(oracle.texture_forward(g["texture"], shifted, r, mode), g["sampled_" + mode])
    n = load("nvdiffrast_64")
    r, d, _ = oracle.rasterize_forward(n["positions"], tri["indices"], (64, 64))
    same_bits(r, n["raster"])
    same_bits(d, n["deriv"])


@pytest.mark.gpu
def test_gpu_matches_fixtures(rasterizer):
    g = load("native_32")
    r, d = rasterizer.forward(g["positions"], g["indices"], (32, 32))
    same_bits(r, g["raster"])
    same_bits(d, g["deriv"])
    gp = rasterizer.backward(g["positions"], g["indices"], r, g["grad_raster"])
    np.testing.assert_allclose(gp, g["grad_pos"], rtol=1e-4, atol=1e-4 * np.abs(g["grad_pos"]).max())
    same_bits(rasterizer.interpolate_forward(g["attrs"], g["indices"], r), g["interp"])
    same_bits(rasterizer.texture_forward(g["texture"], g["uv"], r), g["sampled"])
    gt, gu = rasterizer.texture_backward(g["texture"], g["uv"], r, g["grad_sampled"])
    same_bits(gu, g["grad_uv"])
    np.testing.assert_allclose(gt, g["grad_texture"], rtol=1e-5, atol=1e-7)
    s = load("smoke_16x20")
    r2, _ = rasterizer.forward(g["positions"], g["indices"], (16, 20))
    for mode in ("clamp", "wrap", "mirror"):
        shifted = (s["uv"] + 1.25).astype(np.float32)
        same_bits(rasterizer.texture_forward(s["texture"], shifted, r2, mode), s["sampled_" + mode])
    n = load("nvdiffrast_64")
    r3, d3 = rasterizer.forward(n["positions"], g["indices"], (64, 64))
    same_bits(r3, n["raster"])
    same_bits(d3, n["deriv"])
    gp3 = rasterizer.backward(n["positions"], g["indices"], r3, n["grad_raster"])
    np.testing.assert_allclose(gp3, n["grad_pos"], rtol=1e-4, atol=1e-4 * np.abs(n["grad_pos"]).max())
