"""Differentiable affine Lanczos3 warp and photometric registration refinement.

``affine_lanczos3(image, source_map, fill=0.0)`` is a ``torch.autograd.Function``
over the CPU-path Lanczos3 of src/glass/engine/warp.py:121-228 for an affine
source map s = B (x, y, 1) (B: 6 parameters, reference pixel -> source pixel).
Its backward is a hand-written HIP adjoint (libglass_hip.so) with respect to
both the image and the 6 parameters; stencils are recomputed per pixel, not
stored.  ``refine_affine`` runs Gauss-Newton on the photometric residual of a
resident frame pair natively (the analogue of the reference's translation
refinement, native_bindings.cpp:10079-10416, extended to the full affine).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N

_DP = ctypes.POINTER(ctypes.c_double)


def _map_host(source_map) -> np.ndarray:
    if type(source_map).__module__.startswith("torch"):
        source_map = source_map.detach().to("cpu")
        source_map = source_map.numpy()
    m = np.ascontiguousarray(np.asarray(source_map, dtype=np.float64).reshape(-1))
    if m.size != 6:
        raise ValueError("source_map must hold 6 parameters (2x3)")
    return m


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_image(image):
    if not (type(image).__module__.startswith("torch") and image.is_cuda):
        raise ValueError("image must be a GPU tensor")
    if image.dim() != 2 or image.dtype.itemsize != 4 or not image.is_floating_point():
        raise ValueError("image must be a 2-D float32 tensor")


def _build():
    import torch

    class AffineLanczos3(torch.autograd.Function):
        @staticmethod
        def forward(ctx, image, source_map, fill=0.0):
            _check_image(image)
            img = image.contiguous()
            m = _map_host(source_map)
            out = torch.empty_like(img)
            h, w = img.shape
            N.call("gh_affine_lanczos3_forward", ctypes.c_void_p(img.data_ptr()), h, w, m.ctypes.data_as(_DP),
                   float(fill), ctypes.c_void_p(out.data_ptr()), _stream())
            ctx.save_for_backward(img)
            ctx.map_host = m
            ctx.map_meta = (source_map.dtype, source_map.device) if isinstance(source_map, torch.Tensor) else None
            return out

        @staticmethod
        def backward(ctx, grad_out):
            (img,) = ctx.saved_tensors
            g = grad_out.contiguous().to(torch.float32)
            h, w = img.shape
            grad_img = torch.zeros_like(img) if ctx.needs_input_grad[0] else None
            grad_map = None
            ws = None
            if ctx.needs_input_grad[1]:
                n_ws = ctypes.c_int64(0)
                N.call("gh_affine_workspace_size", h, w, ctypes.byref(n_ws))
                ws = torch.empty(int(n_ws.value), dtype=torch.float64, device=img.device)
                grad_map = torch.empty(6, dtype=torch.float64, device=img.device)
            N.call("gh_affine_lanczos3_backward", ctypes.c_void_p(img.data_ptr()), h, w,
                   ctx.map_host.ctypes.data_as(_DP), ctypes.c_void_p(g.data_ptr()),
                   ctypes.c_void_p(grad_img.data_ptr()) if grad_img is not None else None,
                   ctypes.c_void_p(grad_map.data_ptr()) if grad_map is not None else None,
                   ctypes.c_void_p(ws.data_ptr()) if ws is not None else None,
                   int(ws.numel()) if ws is not None else 0, _stream())
            if grad_map is not None and ctx.map_meta is not None:
                dtype, device = ctx.map_meta
                grad_map = grad_map.to(device=device, dtype=dtype)
            return grad_img, grad_map, None

    return AffineLanczos3


_FUNCTION = None


def affine_lanczos3(image, source_map, fill: float = 0.0):
    """Differentiable warp: out(x, y) = Lanczos3(image) at B (x, y, 1)."""
    global _FUNCTION
    if _FUNCTION is None:
        _FUNCTION = _build()
    return _FUNCTION.apply(image, source_map, fill)


def source_map_of(matrix) -> np.ndarray:
    """The 6 source-map parameters (reference pixel -> moving pixel) of a
    moving -> reference affine matrix."""
    from .resident import invert_affine

    return np.ascontiguousarray(invert_affine(np.asarray(matrix, dtype=np.float64))[:2].reshape(-1))


def refine_affine(stack, reference_index: int, moving_index: int, matrix, max_iterations: int = 20,
                  tolerance_px: float = 1.0e-4):
    """Gauss-Newton refinement of the moving -> reference affine matrix of a
    resident frame pair.  Returns (3x3 matrix, report dict)."""
    m = np.ascontiguousarray(np.asarray(matrix, dtype=np.float64).reshape(9))
    out = (ctypes.c_double * 9)()
    rep = (ctypes.c_double * 4)()
    N.call("gh_refine_affine", stack._h, int(reference_index), int(moving_index), m.ctypes.data_as(_DP),
           int(max_iterations), float(tolerance_px), out, rep)
    return np.asarray(list(out), dtype=np.float64).reshape(3, 3), {
        "iterations": int(rep[0]), "rms": float(rep[1]), "pixels": int(rep[2]), "last_step_px": float(rep[3])}
