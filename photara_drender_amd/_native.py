"""ctypes binding of ``libpdr.so`` (the C ABI of ``include/pdr.h``).

Used by the zero-copy PyTorch path (device pointers, torch's current stream)
and by the benchmark.  Loading fails loudly when the library is missing: there
is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# PDR_LIB names another build of the same library (the host-sanitizer build)
LIB_PATH = os.environ.get("PDR_LIB") or os.path.join(_HERE, "libpdr.so")

_F = ctypes.c_float
_U32 = ctypes.c_uint32
_P = ctypes.c_void_p
_I = ctypes.c_int

PDR_FLAG_VALIDATE_DEVICE_INPUTS = 1
PDR_FLAG_DETERMINISTIC = 2
PDR_FLAG_AUTOTUNE = 4

CULL_MODES = {"none": 0, "back": 1, "front": 2}
ADDRESS_MODES = {"clamp": 0, "wrap": 1, "mirror": 2}

_ERRORS = {1: ValueError, 2: IndexError, 3: OverflowError, 4: RuntimeError}


class RasterView(ctypes.Structure):
    _fields_ = [("positions", _P), ("width", _U32), ("height", _U32), ("viewport", _F * 4), ("raster", _P),
                ("derivatives", _P), ("grad_raster", _P), ("grad_positions", _P)]


class InterpView(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("raster", _P), ("out", _P), ("grad", _P), ("grad_raster", _P)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("vendor_id", _U32), ("device_id", _U32), ("api_version", _U32),
                ("supports_ray_query", _U32)]


class RefineView(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("channels", _U32), ("image", _P), ("visibility_mask", _P),
                ("world_to_clip", _F * 16), ("has_viewport", _I), ("viewport", _F * 4)]


class RefineOptions(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("steps", _U32), ("batch_size", _U32), ("learning_rate", _F),
                ("minimum_learning_rate", _F), ("photometric_epsilon", _F), ("adam_beta1", _F), ("adam_beta2", _F),
                ("adam_epsilon", _F), ("clamp_min", _F), ("clamp_max", _F), ("seam_polish_steps", _U32),
                ("seam_learning_rate", _F), ("seam_epsilon", _F), ("tv_weight", _F), ("tv_epsilon", _F),
                ("chart_ids", _P), ("joint_seam_weight", _F), ("use_graph", _U32), ("checkpoint_every", _U32),
                ("checkpoints", _P), ("stop_after", _U32)]


class BakeView(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("channels", _U32), ("image", _P), ("visibility_mask", _P),
                ("world_to_clip", _F * 16), ("camera_position", _F * 3), ("has_viewport", _I), ("viewport", _F * 4),
                ("weight", _F), ("depth_bias", _F), ("min_view_cosine", _F)]


class BakeOptions(ctypes.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("blend_mode", _U32), ("visibility_mode", _U32),
                ("pcf_radius", _U32), ("ray_origin_bias", _F), ("allow_visibility_fallback", _I), ("mask_floor", _F),
                ("softmax_scale", _F)]


class PaddingOptions(ctypes.Structure):
    _fields_ = [("margin", _U32), ("fill_unobserved", _I), ("maximum_fill_distance", _U32), ("valid_threshold", _F)]


class PaddingStats(ctypes.Structure):
    _fields_ = [("valid_texels", ctypes.c_uint64), ("gutter_texels", ctypes.c_uint64),
                ("unobserved_texels", ctypes.c_uint64), ("remaining_unobserved", ctypes.c_uint64),
                ("filled_texels", ctypes.c_uint64), ("seconds", ctypes.c_double)]


_SIGNATURES = {
    "pdr_last_error": (ctypes.c_char_p, []),
    "pdr_context_create": (_I, [_U32, _I, ctypes.POINTER(_P)]),
    "pdr_context_destroy": (None, [_P]),
    "pdr_context_set_flags": (_I, [_P, _U32]),
    "pdr_context_flags": (_U32, [_P]),
    "pdr_context_device_info": (_I, [_P, ctypes.POINTER(DeviceInfo)]),
    "pdr_enumerate_devices": (_I, [ctypes.POINTER(DeviceInfo), _U32, ctypes.POINTER(_U32)]),
    "pdr_synchronize": (_I, [_P, _P]),
    "pdr_autotune_export": (_I, [_P, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "pdr_autotune_import": (_I, [_P, ctypes.c_char_p]),
    "pdr_autotune_clear": (_I, [_P]),
    "pdr_context_set_timing": (_I, [_P, _I]),
    "pdr_context_kernel_time": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_uint64)]),
    "pdr_rasterize_forward": (_I, [_P, _P, _U32, _P, _U32, _U32, _U32, _U32, _I, _P, _P, _P, _P, _P]),
    "pdr_rasterize_forward_noalloc": (_I, [_P, _P, _U32, _P, _U32, _U32, _U32, _U32, _I, _P, _P, _P, _P, _P]),
    "pdr_rasterize_forward_batched": (_I, [_P, ctypes.POINTER(RasterView), _U32, _U32, _P, _U32, _U32, _I, _P]),
    "pdr_rasterize_backward": (_I, [_P, _P, _U32, _P, _U32, _P, _U32, _U32, _P, _P, _P, _P]),
    "pdr_rasterize_backward_batched": (_I, [_P, ctypes.POINTER(RasterView), _U32, _U32, _P, _U32, _P]),
    "pdr_interpolate_forward": (_I, [_P, _P, _U32, _U32, _P, _U32, _P, _U32, _U32, _P, _P]),
    "pdr_interpolate_backward": (_I, [_P, _P, _U32, _U32, _P, _U32, _P, _U32, _U32, _P, _P, _P, _P]),
    "pdr_interpolate_forward_batched": (_I, [_P, ctypes.POINTER(InterpView), _U32, _P, _U32, _U32, _P, _U32, _P]),
    "pdr_interpolate_backward_batched": (_I, [_P, ctypes.POINTER(InterpView), _U32, _P, _U32, _U32, _P, _U32, _P,
                                              _P]),
    "pdr_texture_forward": (_I, [_P, _P, _U32, _U32, _U32, _U32, _P, _P, _U32, _U32, _P, _P]),
    "pdr_texture_backward": (_I, [_P, _P, _U32, _U32, _U32, _U32, _P, _P, _U32, _U32, _P, _I, _P, _P, _P]),
    "pdr_refine_options_default": (None, [ctypes.POINTER(RefineOptions)]),
    "pdr_bake_options_default": (None, [ctypes.POINTER(BakeOptions)]),
    "pdr_bake_view_default": (None, [ctypes.POINTER(BakeView)]),
    "pdr_suggest_softmax_scale": (_F, [_P, _U32, _P, _U32]),
    "pdr_bake": (_I, [_P, _P, _P, _P, _U32, _P, _U32, ctypes.POINTER(BakeView), _U32, ctypes.POINTER(BakeOptions), _P,
                      _P, _P, _P, _P, _P, ctypes.POINTER(_I)]),
    "pdr_refine": (_I, [_P, _P, _P, _U32, _P, _P, _U32, ctypes.POINTER(RefineView), _U32, _P, _U32,
                        ctypes.POINTER(RefineOptions), _P, _P, _P, ctypes.POINTER(ctypes.c_double)]),
    "pdr_padding_options_default": (None, [ctypes.POINTER(PaddingOptions)]),
    "pdr_pad_texture_atlas": (_I, [_P, _P, _U32, _U32, _U32, _P, _P, _P, ctypes.POINTER(PaddingOptions),
                                   ctypes.POINTER(PaddingStats), _P]),
    "pdr_bake_features": (_I, [_P, _P, _P, _P, _U32, _P, _U32, ctypes.POINTER(BakeView), _U32,
                               ctypes.POINTER(BakeOptions), _P]),
    "pdr_mlp_infer": (_I, [_P, _P, ctypes.c_uint64, _P, _P, _P]),
    "pdr_mlp_gradient": (_I, [_P, _P, _P, _P, ctypes.c_uint64, _P, _P, _P, _P]),
    "pdr_mlp_adam_step": (_I, [_P, _P, _P, _P, ctypes.c_uint64, _P, _P, _P, _U32, _F, _F, _F, _F, _P, _P]),
}

MLP_FEATURES = 40
MLP_PARAMS = 6979

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

# Host-only self-test hooks exported by libpdr outside the public header.
_TESTING_SIGNATURES = {
    "pdr_testing_candidate_region": (_I, [_P, _U32, _U32, _P, _U32, _I, _P, _P]),
    "pdr_testing_tv_gradient": (_I, [_P, _U32, _U32, _P, _F, _F, _P]),
    "pdr_testing_set_pad_poll_limit": (None, [_U32]),
    "pdr_testing_pad_fallbacks": (ctypes.c_uint64, []),
}

_lib = None


def lib():
    """The loaded libpdr (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        try:  # bind libpdr to torch's HIP runtime when torch is present (one runtime per process)
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise OSError(f"libpdr.so not found at {LIB_PATH}; run __graft_entry__.build() first")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in {**_SIGNATURES, **_TESTING_SIGNATURES}.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status:
        raise _ERRORS.get(status, RuntimeError)(lib().pdr_last_error().decode())


class Context:
    """Owning handle of a pdr_context (device index, stream, scratch)."""

    def __init__(self, device_index: int = 0, flags: int | None = None):
        h = _P()
        check(lib().pdr_context_create(device_index, 0, ctypes.byref(h)))
        self.handle = h.value
        if flags is not None:
            check(lib().pdr_context_set_flags(self.handle, flags))

    def close(self):
        if getattr(self, "handle", None):
            lib().pdr_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_flags(self, flags: int) -> None:
        check(lib().pdr_context_set_flags(self.handle, flags))

    def set_timing(self, enable: bool) -> None:
        check(lib().pdr_context_set_timing(self.handle, int(bool(enable))))

    def kernel_time(self, name: str):
        """(total milliseconds, launches) of one kernel since timing was enabled."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        check(lib().pdr_context_kernel_time(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    @property
    def flags(self) -> int:
        return int(lib().pdr_context_flags(self.handle))

    def autotune_export(self) -> str:
        """The autotuner's cache (one workload class per line)."""
        n = ctypes.c_size_t()
        check(lib().pdr_autotune_export(self.handle, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(lib().pdr_autotune_export(self.handle, buf, n.value, ctypes.byref(n)))
        return buf.value.decode()

    def autotune_import(self, text: str) -> None:
        check(lib().pdr_autotune_import(self.handle, text.encode()))

    def autotune_clear(self) -> None:
        check(lib().pdr_autotune_clear(self.handle))
