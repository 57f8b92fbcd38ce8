"""The adjoint's closed form on the CPU: 2 (x - P_M x) (P_M from the oracle's modulus
projection) equals torch's autograd gradient of sum_unmasked (|F x| - A)^2 (FP64)."""

import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("shape", [(16, 12), (8, 6, 10)])
def test_gradient_closed_form_matches_torch_autograd(shape):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    amp = rng.uniform(0.2, 2.0, shape)
    mask = (rng.random(shape) < 0.1).astype(np.uint8)
    # closed form through the oracle (FFT order, unitary transforms, FP64 twin)
    w, num = oracle.modulus_num(oracle.fft(x, True, "d"), amp, mask, "d")
    pm = oracle.fft(w, False, "d")
    grad = 2 * (x - pm)
    # torch reference
    xt = torch.tensor(x, dtype=torch.complex128, requires_grad=True)
    y = torch.fft.fftn(xt, norm="ortho")
    keep = torch.tensor(1 - mask, dtype=torch.float64)
    loss = (((y.abs() - torch.tensor(amp)) ** 2) * keep).sum()
    loss.backward()
    assert abs(loss.item() - num) < 1e-10 * num
    np.testing.assert_allclose(grad, xt.grad.numpy(), rtol=1e-9, atol=1e-11)
