"""PyTorch autograd functions over libpdr (python/photara_drender/torch.py:17-241).

Same functions, arguments, validation messages and gradient routing as the
reference:
* ``rasterize`` -> raster, derivatives; gradients flow to positions only,
  derivatives are non-differentiable;
* ``interpolate`` -> gradients to vertex attributes and to the raster;
* ``texture`` -> gradients to texels and (when requested) to the uv image.

GPU tensors go to the HIP kernels zero-copy: the C ABI receives their device
pointers and torch's current stream, so nothing crosses PCIe.  CPU tensors take
the reference's numpy route through ``_photara_drender`` (staged to the GPU and
back).  Both run the same kernels.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native
from ._photara_drender import Rasterizer


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _u32_indices(indices: torch.Tensor) -> torch.Tensor:
    """Indices as 32-bit words in int32 lanes (the reference casts with astype(uint32))."""
    if indices.dtype == torch.int32:
        return indices.detach().contiguous()
    wrapped = indices.detach().to(torch.int64).bitwise_and(0xFFFFFFFF)
    wrapped = torch.where(wrapped >= 2**31, wrapped - 2**32, wrapped)
    return wrapped.to(torch.int32).contiguous()


def _f32(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float32).contiguous()


def _viewport_arg(viewport):
    if viewport is None:
        return None
    return (ctypes.c_float * 4)(*[float(v) for v in viewport])


class _RasterizeFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, indices, resolution, rasterizer, cull_mode, output_barycentric_derivatives,
                viewport):
        if positions.dtype != torch.float32 or positions.ndim != 2 or positions.shape[1] != 4:
            raise ValueError("positions must be a float32 tensor with shape [vertex_count, 4]")
        if indices.ndim != 2 or indices.shape[1] != 3:
            raise ValueError("indices must have shape [triangle_count, 3]")
        height, width = int(resolution[0]), int(resolution[1])
        if positions.is_cuda:
            if cull_mode not in _native.CULL_MODES:
                raise ValueError("cull_mode must be 'none', 'back', or 'front'")
            pos = positions.detach().contiguous()
            idx = _u32_indices(indices.to(positions.device))
            raster = torch.empty((height, width, 4), dtype=torch.float32, device=positions.device)
            if output_barycentric_derivatives:
                derivatives = torch.empty((height, width, 4), dtype=torch.float32, device=positions.device)
            else:
                derivatives = torch.empty((0, 0, 4), dtype=torch.float32, device=positions.device)
            vp_out = (ctypes.c_float * 4)()
            _native.check(_native.lib().pdr_rasterize_forward(
                rasterizer.context_handle, pos.data_ptr(), pos.shape[0], idx.data_ptr(), idx.shape[0], width, height,
                _native.CULL_MODES[cull_mode], int(bool(output_barycentric_derivatives)), _viewport_arg(viewport),
                raster.data_ptr(), derivatives.data_ptr() if output_barycentric_derivatives else None, vp_out,
                _stream()))
        else:
            positions_numpy = positions.detach().cpu().contiguous().numpy()
            indices_numpy = indices.detach().cpu().contiguous().numpy().astype(np.uint32, copy=False)
            raster_numpy, derivative_numpy = rasterizer.forward(positions_numpy, indices_numpy, (height, width),
                                                                cull_mode, output_barycentric_derivatives, viewport)
            raster = torch.from_numpy(raster_numpy)
            derivatives = torch.from_numpy(derivative_numpy)
        ctx.rasterizer = rasterizer
        ctx.viewport = viewport
        ctx.save_for_backward(positions.detach(), indices.detach(), raster.detach())
        ctx.mark_non_differentiable(derivatives)
        return raster, derivatives

    @staticmethod
    def backward(ctx, grad_raster, _grad_derivatives):
        positions, indices, raster = ctx.saved_tensors
        if grad_raster is None:
            return None, None, None, None, None, None, None
        if positions.is_cuda:
            pos = positions.contiguous()
            idx = _u32_indices(indices.to(positions.device))
            g = _f32(grad_raster)
            r = raster.contiguous()
            out = torch.empty((pos.shape[0], 4), dtype=torch.float32, device=pos.device)
            _native.check(_native.lib().pdr_rasterize_backward(
                ctx.rasterizer.context_handle, pos.data_ptr(), pos.shape[0], idx.data_ptr(), idx.shape[0],
                r.data_ptr(), r.shape[1], r.shape[0], g.data_ptr(), _viewport_arg(ctx.viewport), out.data_ptr(),
                _stream()))
            return out, None, None, None, None, None, None
        grad_positions = ctx.rasterizer.backward(
            positions.cpu().contiguous().numpy(),
            indices.cpu().contiguous().numpy().astype(np.uint32, copy=False),
            raster.cpu().contiguous().numpy(),
            grad_raster.detach().cpu().contiguous().numpy().astype(np.float32, copy=False),
            ctx.viewport,
        )
        return torch.from_numpy(grad_positions), None, None, None, None, None, None


class _InterpolateFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertex_attributes, raster, indices, rasterizer):
        if vertex_attributes.dtype != torch.float32 or vertex_attributes.ndim != 2:
            raise ValueError("vertex_attributes must be a float32 tensor with shape [vertex_count, attribute_count]")
        if raster.dtype != torch.float32 or raster.ndim != 3 or raster.shape[2] != 4:
            raise ValueError("raster must be a float32 tensor with shape [height, width, 4]")
        if indices.ndim != 2 or indices.shape[1] != 3:
            raise ValueError("indices must have shape [triangle_count, 3]")
        if vertex_attributes.device != raster.device:
            raise ValueError("vertex_attributes and raster must be on the same device")
        if vertex_attributes.is_cuda:
            attr = vertex_attributes.detach().contiguous()
            r = raster.detach().contiguous()
            idx = _u32_indices(indices.to(attr.device))
            h, w = r.shape[:2]
            out = torch.empty((h, w, attr.shape[1]), dtype=torch.float32, device=attr.device)
            _native.check(_native.lib().pdr_interpolate_forward(
                rasterizer.context_handle, attr.data_ptr(), attr.shape[0], attr.shape[1], idx.data_ptr(),
                idx.shape[0], r.data_ptr(), w, h, out.data_ptr(), _stream()))
        else:
            out = torch.from_numpy(rasterizer.interpolate_forward(
                vertex_attributes.detach().cpu().contiguous().numpy(),
                indices.detach().cpu().contiguous().numpy().astype(np.uint32, copy=False),
                raster.detach().cpu().contiguous().numpy()))
        ctx.rasterizer = rasterizer
        ctx.save_for_backward(vertex_attributes.detach(), raster.detach(), indices.detach())
        return out

    @staticmethod
    def backward(ctx, grad_interpolated):
        vertex_attributes, raster, indices = ctx.saved_tensors
        if vertex_attributes.is_cuda:
            attr = vertex_attributes.contiguous()
            r = raster.contiguous()
            idx = _u32_indices(indices.to(attr.device))
            g = _f32(grad_interpolated)
            h, w = r.shape[:2]
            ga = torch.empty_like(attr)
            gr = torch.empty((h, w, 4), dtype=torch.float32, device=attr.device)
            _native.check(_native.lib().pdr_interpolate_backward(
                ctx.rasterizer.context_handle, attr.data_ptr(), attr.shape[0], attr.shape[1], idx.data_ptr(),
                idx.shape[0], r.data_ptr(), w, h, g.data_ptr(), ga.data_ptr(), gr.data_ptr(), _stream()))
            return ga, gr, None, None
        ga, gr = ctx.rasterizer.interpolate_backward(
            vertex_attributes.cpu().contiguous().numpy(),
            indices.cpu().contiguous().numpy().astype(np.uint32, copy=False),
            raster.cpu().contiguous().numpy(),
            grad_interpolated.detach().cpu().contiguous().numpy().astype(np.float32, copy=False),
        )
        return torch.from_numpy(ga), torch.from_numpy(gr), None, None


class _TextureFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, texture_values, uv, raster, rasterizer, address_mode):
        if texture_values.dtype != torch.float32 or texture_values.ndim != 3:
            raise ValueError("texture_values must be float32 with shape [texture_height, texture_width, channels]")
        if uv.dtype != torch.float32 or uv.ndim != 3 or uv.shape[2] != 2:
            raise ValueError("uv must be float32 with shape [image_height, image_width, 2]")
        if raster.dtype != torch.float32 or raster.ndim != 3 or raster.shape[2] != 4:
            raise ValueError("raster must be float32 with shape [image_height, image_width, 4]")
        if uv.shape[:2] != raster.shape[:2]:
            raise ValueError("uv and raster image dimensions must match")
        if texture_values.device != uv.device or uv.device != raster.device:
            raise ValueError("texture_values, uv, and raster must be on the same device")
        if texture_values.is_cuda:
            if address_mode not in _native.ADDRESS_MODES:
                raise ValueError("address_mode must be 'clamp', 'wrap', or 'mirror'")
            tex = texture_values.detach().contiguous()
            u = uv.detach().contiguous()
            r = raster.detach().contiguous()
            th, tw, c = tex.shape
            h, w = r.shape[:2]
            out = torch.empty((h, w, c), dtype=torch.float32, device=tex.device)
            _native.check(_native.lib().pdr_texture_forward(
                rasterizer.context_handle, tex.data_ptr(), tw, th, c, _native.ADDRESS_MODES[address_mode],
                u.data_ptr(), r.data_ptr(), w, h, out.data_ptr(), _stream()))
        else:
            out = torch.from_numpy(rasterizer.texture_forward(
                texture_values.detach().cpu().contiguous().numpy(), uv.detach().cpu().contiguous().numpy(),
                raster.detach().cpu().contiguous().numpy(), address_mode))
        ctx.rasterizer = rasterizer
        ctx.address_mode = address_mode
        ctx.save_for_backward(texture_values.detach(), uv.detach(), raster.detach())
        return out

    @staticmethod
    def backward(ctx, grad_sampled):
        texture_values, uv, raster = ctx.saved_tensors
        compute_uv_gradient = bool(ctx.needs_input_grad[1])
        if texture_values.is_cuda:
            tex = texture_values.contiguous()
            u = uv.contiguous()
            r = raster.contiguous()
            g = _f32(grad_sampled)
            th, tw, c = tex.shape
            h, w = r.shape[:2]
            gt = torch.empty_like(tex)
            gu = torch.empty((h, w, 2), dtype=torch.float32, device=tex.device) if compute_uv_gradient else None
            _native.check(_native.lib().pdr_texture_backward(
                ctx.rasterizer.context_handle, tex.data_ptr(), tw, th, c, _native.ADDRESS_MODES[ctx.address_mode],
                u.data_ptr(), r.data_ptr(), w, h, g.data_ptr(), int(compute_uv_gradient), gt.data_ptr(),
                gu.data_ptr() if gu is not None else None, _stream()))
            return gt, gu, None, None, None
        gt, gu = ctx.rasterizer.texture_backward(
            texture_values.cpu().contiguous().numpy(), uv.cpu().contiguous().numpy(),
            raster.cpu().contiguous().numpy(),
            grad_sampled.detach().cpu().contiguous().numpy().astype(np.float32, copy=False),
            ctx.address_mode, compute_uv_gradient)
        return torch.from_numpy(gt), None if gu is None else torch.from_numpy(gu), None, None, None


def rasterize(
    positions: torch.Tensor,
    indices: torch.Tensor,
    resolution: Sequence[int],
    *,
    rasterizer: Optional[Rasterizer] = None,
    device_index: int = 0,
    enable_validation: bool = False,
    cull_mode: str = "none",
    output_barycentric_derivatives: bool = True,
    viewport: Optional[Sequence[float]] = None,
) -> Tuple[torch.Tensor, torch.Tensor]:
    """Rasterize clip-space vertices (resolution is (height, width)) and connect the output to autograd."""
    if len(resolution) != 2 or int(resolution[0]) <= 0 or int(resolution[1]) <= 0:
        raise ValueError("resolution must be (height, width) with positive values")
    active = rasterizer or Rasterizer(device_index, enable_validation)
    vp = None
    if viewport is not None:
        if len(viewport) != 4:
            raise ValueError("viewport must be (x, y, width, height)")
        vp = tuple(float(v) for v in viewport)
    return _RasterizeFunction.apply(positions, indices, (int(resolution[0]), int(resolution[1])), active, cull_mode,
                                    output_barycentric_derivatives, vp)


def interpolate(vertex_attributes: torch.Tensor, raster: torch.Tensor, indices: torch.Tensor, *,
                rasterizer: Rasterizer) -> torch.Tensor:
    """Interpolate vertex attributes; gradients reach the attributes and the raster."""
    return _InterpolateFunction.apply(vertex_attributes, raster, indices, rasterizer)


def texture(texture_values: torch.Tensor, uv: torch.Tensor, raster: torch.Tensor, *, rasterizer: Rasterizer,
            address_mode: str = "clamp") -> torch.Tensor:
    """Bilinear texel-centre sampling; gradients reach the texels and the uv image."""
    return _TextureFunction.apply(texture_values, uv, raster, rasterizer, address_mode)
