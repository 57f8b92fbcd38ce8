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
    (float64, shape (batch,)).  noalloc: call ``qpch_amplitude_loss_noalloc`` (scratch in the
    library's per-GPU array, nothing allocated: the form a CUDA-graph capture needs; calls
    on one GPU must not overlap)."""

    @staticmethod
    def forward(ctx, x, amp, noalloc=False):
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
        call = _qpch.amplitude_loss_noalloc if noalloc else _qpch.amplitude_loss
        call(xc.data_ptr(), a.data_ptr(), 0 if shared else n, grad.data_ptr(), loss.data_ptr(), shape, x.shape[0],
             precision, _stream(x.device))
        ctx.save_for_backward(grad)
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        (grad,) = ctx.saved_tensors
        g = grad_loss.to(grad.real.dtype).reshape(-1, *([1] * (grad.dim() - 1)))
        return grad * g, None, None


def amplitude_loss(x: torch.Tensor, amplitude_fft_order: torch.Tensor, noalloc: bool = False) -> torch.Tensor:
    """Per-object amplitude losses of x (batch, *shape), differentiable in x."""
    return AmplitudeLoss.apply(x, amplitude_fft_order, noalloc)


def refine(x0, intensity, support, mask=None, steps=500, lr=1e-2, device="cuda", graph=False, capturable=None,
           fused=None):
    """Adam on the support-projected iterate: minimises L(P_S z) from z = x0.

    x0: (batch, *shape) complex (centre-origin); intensity: centre-origin, shared or per
    object; support: bool, shared or per object.  Returns (P_S z, loss history (steps, batch)).

    graph=True runs three steps eagerly, then captures one whole step (the loss through the
    no-allocation entry point, its adjoint, the Adam update with its state on the device, and
    the history row it writes) in a CUDA graph and replays it for the remaining steps: the
    same kernels in the same order, without a host round of autograd and optimizer calls per
    step.  capturable (default: graph) selects Adam's on-device step count and bias
    correction, which a graph needs; with the same setting the two forms give the same bits.
    fused: torch's single-kernel Adam (default: on for the graph form, whose replays it cuts
    from about twenty kernels to a few)."""
    x0 = torch.as_tensor(np.asarray(x0), device=device)
    if not x0.is_complex():
        x0 = x0.to(torch.complex64)
    ndim = x0.dim() - 1
    amp = _to_fft_order(signed_amplitude(intensity, mask, device=device), ndim)
    s = torch.as_tensor(np.asarray(support, dtype=bool), device=device).to(x0.real.dtype)
    # The parameter is the real view of z (re, im pairs): Adam on a complex parameter
    # updates exactly those elements, and the fused optimizer kernel takes real tensors.
    zr = torch.view_as_real(x0.clone()).clone().requires_grad_(True)
    cap = graph if capturable is None else capturable
    opt = torch.optim.Adam([zr], lr=lr, capturable=cap, fused=graph if fused is None else fused)
    history = torch.empty((steps, x0.shape[0]), dtype=torch.float64, device=device)

    def step():
        loss = amplitude_loss(torch.view_as_complex(zr) * s, amp, noalloc=graph)
        loss.sum().backward()
        opt.step()
        return loss.detach()

    if not graph:
        for k in range(steps):
            opt.zero_grad(set_to_none=True)
            history[k] = step()
    else:
        # eager warm-up on a side stream (Adam's state and the library's scratch address come
        # into being here, before any capture); these are real steps of the run
        warm = min(steps, 3)
        side = torch.cuda.Stream(device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            for k in range(warm):
                opt.zero_grad(set_to_none=True)
                history[k] = step()
        torch.cuda.current_stream(device).wait_stream(side)
        if steps > warm:
            slot = torch.full((1,), warm, dtype=torch.long, device=device)  # history row of the next step
            opt.zero_grad(set_to_none=True)  # the captured backward writes the gradient afresh each replay
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                history.index_copy_(0, slot, step()[None])
                slot.add_(1)
            for _ in range(steps - warm):
                g.replay()
    with torch.no_grad():
        out = torch.view_as_complex(zr) * s
    return out, history
