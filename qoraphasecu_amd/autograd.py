"""Differentiable amplitude loss on the GPU, for gradient-based refinement after the
projection recipe (no counterpart in the reference).

    L(x) = sum over usable pixels of (|F x| - sqrt(I))^2      (F the unitary DFT)

The forward and its hand-written adjoint run in one call of the C ABI
(``qpch_amplitude_loss``).  That call reuses the projection kernels: a forward transform, the
Fourier-modulus projection with the FP64 loss sum, an inverse transform, and the gradient
``2 (x - P_M x)``.  ``AmplitudeLoss`` is a ``torch.autograd.Function`` whose backward only
scales the saved gradient.

Arrays follow the package convention: objects and intensities are centre-origin.  The
loss does not depend on the real-space origin, so only the amplitude is reordered, once.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _qpch


def _stream(device) -> int:
    return int(torch.cuda.current_stream(device).cuda_stream)


def signed_amplitude(intensity, mask=None, dtype=torch.float32, device="cuda") -> torch.Tensor:
    """sqrt(max(I, 0)) with masked pixels (mask = True) set to -1, still centre-origin.

    ``intensity`` is shaped like one object or stacked as (batch, *shape) for one
    amplitude per object."""
    inten = np.asarray(intensity, dtype=np.float64)
    amp = np.sqrt(np.maximum(inten, 0.0)).astype(np.float32)  # the float amplitude of make_problem
    if mask is not None:
        amp = np.where(np.asarray(mask, dtype=bool), np.float32(-1.0), amp)
    return torch.as_tensor(amp, device=device).to(dtype)


def _to_fft_order(a: torch.Tensor, ndim: int) -> torch.Tensor:
    return torch.fft.ifftshift(a, dim=tuple(range(a.dim() - ndim, a.dim())))


def fft_order_amplitude(intensity, mask=None, ndim=None, dtype=torch.float32, device="cuda") -> torch.Tensor:
    """The signed amplitude AmplitudeLoss takes: sqrt(max(I, 0)), -1 where masked, FFT order."""
    a = signed_amplitude(intensity, mask, dtype, device)
    return _to_fft_order(a, ndim or a.dim())


class AmplitudeLoss(torch.autograd.Function):
    """x: complex tensor (batch, *shape) on the GPU; amp: signed amplitude in FFT order,
    shaped like one object (shared) or (batch, *shape).  Returns the per-object losses
    (float64, shape (batch,))."""

    @staticmethod
    def forward(ctx, x, amp):
        if not x.is_cuda or not x.is_complex():
            raise ValueError("x must be a complex tensor on the GPU")
        if x.dtype not in (torch.complex64, torch.complex128):
            raise ValueError("x must be complex64 or complex128")
        precision = 0 if x.dtype == torch.complex64 else 1
        real = torch.float32 if precision == 0 else torch.float64
        ndim = amp.dim() if amp.shape == x.shape[1:] else amp.dim() - 1
        shape = tuple(x.shape[1:])
        if ndim not in (2, 3) or tuple(amp.shape[-ndim:]) != shape:
            raise ValueError(f"amplitude shape {tuple(amp.shape)} does not match objects {tuple(x.shape)}")
        shared = amp.shape == x.shape[1:]
        if not shared and amp.shape[0] != x.shape[0]:
            raise ValueError("a per-object amplitude needs one array per object")
        xc = x.detach().contiguous()
        a = amp.detach().to(device=x.device, dtype=real).contiguous()
        grad = torch.empty_like(xc)
        loss = torch.empty(x.shape[0], dtype=torch.float64, device=x.device)
        n = int(np.prod(shape))
        _qpch.amplitude_loss(xc.data_ptr(), a.data_ptr(), 0 if shared else n, grad.data_ptr(), loss.data_ptr(),
                             shape, x.shape[0], precision, _stream(x.device))
        ctx.save_for_backward(grad)
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        (grad,) = ctx.saved_tensors
        g = grad_loss.to(grad.real.dtype).reshape(-1, *([1] * (grad.dim() - 1)))
        return grad * g, None


def amplitude_loss(x: torch.Tensor, amplitude_fft_order: torch.Tensor) -> torch.Tensor:
    """Per-object amplitude losses of x (batch, *shape), differentiable in x."""
    return AmplitudeLoss.apply(x, amplitude_fft_order)


def refine(x0, intensity, support, mask=None, steps=500, lr=1e-2, device="cuda"):
    """Adam on the support-projected iterate: minimises L(P_S z) from z = x0.

    x0: (batch, *shape) complex (centre-origin); intensity: centre-origin, shared or per
    object; support: bool, shared or per object.  Returns (P_S z, loss history (steps, batch))."""
    x0 = torch.as_tensor(np.asarray(x0), device=device)
    if not x0.is_complex():
        x0 = x0.to(torch.complex64)
    ndim = x0.dim() - 1
    amp = _to_fft_order(signed_amplitude(intensity, mask, device=device), ndim)
    s = torch.as_tensor(np.asarray(support, dtype=bool), device=device).to(x0.real.dtype)
    z = x0.clone().requires_grad_(True)
    opt = torch.optim.Adam([z], lr=lr)
    history = []
    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = amplitude_loss(z * s, amp)
        loss.sum().backward()
        opt.step()
        history.append(loss.detach())
    with torch.no_grad():
        out = z * s
    return out, torch.stack(history)
