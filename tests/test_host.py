"""CPU-side checks of the product: the C-ABI library loads and exports every
symbol include/pdr.h declares; the ctypes table matches the header; the Python
package imports its native module; workload generators have the sizes the
benchmark configurations name."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pdr.h")).read()
    return sorted(set(re.findall(r"\b(pdr_[a-z_0-9]+)\s*\(", text)))


def test_libpdr_exports_every_declared_symbol():
    from photara_drender_amd import _native

    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_ctypes_table_covers_header():
    from photara_drender_amd import _native

    assert set(declared_symbols()) == set(_native.EXPORTED_SYMBOLS)


def test_native_module_imports_and_has_reference_surface():
    import photara_drender_amd as pdr

    for name in ("Rasterizer", "TextureBaker", "DeviceInfo", "enumerate_devices", "rasterize", "interpolate",
                 "texture", "render_textured_mesh", "masked_charbonnier_loss", "atlas_total_variation",
                 "seam_consistency_loss", "optimize_texture_atlas", "AtlasOptimizationOptions"):
        assert getattr(pdr, name) is not None, name
    for meth in ("forward", "backward", "interpolate_forward", "interpolate_backward", "texture_forward",
                 "texture_backward"):
        assert hasattr(pdr.Rasterizer, meth)
    assert hasattr(pdr.TextureBaker, "refine_texture")


def test_context_without_gpu_raises_runtime_error():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import photara_drender_amd as pdr

    with pytest.raises(RuntimeError):
        pdr.Rasterizer()


def test_refine_options_default_matches_reference_types():
    from photara_drender_amd import _native

    o = _native.RefineOptions()
    _native.lib().pdr_refine_options_default(ctypes.byref(o))
    # include/photara_drender/types.hpp:267-283
    assert (o.width, o.height, o.steps, o.batch_size) == (1024, 1024, 1000, 4)
    assert o.learning_rate == np.float32(5e-3) and o.minimum_learning_rate == np.float32(2.5e-4)
    assert o.photometric_epsilon == np.float32(1e-3)
    assert (o.adam_beta1, o.adam_beta2) == (np.float32(0.9), np.float32(0.999))
    assert o.adam_epsilon == np.float32(1e-8) and (o.clamp_min, o.clamp_max) == (0.0, 1.0)
    assert o.seam_polish_steps == 30 and o.seam_learning_rate == np.float32(1e-3)
    assert o.seam_epsilon == np.float32(1e-4)
    assert o.stop_after == 0 and o.checkpoint_every == 0 and o.use_graph == 1  # extensions off by default


def test_workload_sizes():
    from photara_drender_amd import scenes

    pos, idx = scenes.icosphere(5)
    assert (pos.shape[0], idx.shape[0]) == (10242, 20480)
    pos, idx = scenes.heightfield()
    assert (pos.shape[0], idx.shape[0]) == (100651, 200000)
    _, _, idx = scenes.uv_sphere(32, 33)
    assert idx.shape[0] == 2048
    sizes = scenes.MULTIVIEW_SIZES * 2
    assert sum(h * w for h, w in sizes) == 15_163_392
