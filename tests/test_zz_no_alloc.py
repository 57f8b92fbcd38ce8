"""qpch_amplitude_loss_noalloc (the entry point that allocates nothing) against
qpch_amplitude_loss through the C ABI: the same losses and gradients, bit for bit (they run
the same kernels), on two calls in a row with the second one smaller, for the fused fp32
passes (2D and 3D), the generic fp32 kernels (extents that are not powers of two) and the
FP64 path; and a call needing more scratch than the array holds is refused before any work
is queued."""

import pytest
import torch

from qoraphasecu_amd import _qpch

pytestmark = pytest.mark.gpu


def _inputs(shape, batch, dtype, seed):
    g = torch.Generator().manual_seed(seed)
    real = torch.float64 if dtype == torch.complex128 else torch.float32
    x = torch.complex(torch.randn((batch, *shape), generator=g, dtype=real),
                      torch.randn((batch, *shape), generator=g, dtype=real))
    amp = torch.rand(shape, generator=g, dtype=real) * 3.0
    amp.view(-1)[::97] = -1.0  # masked pixels
    return x.cuda(), amp.cuda()


def _call(entry, x, amp, shape, batch, precision):
    grad = torch.full_like(x, float("nan"))
    loss = torch.full((batch,), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    entry(x.data_ptr(), amp.data_ptr(), 0, grad.data_ptr(), loss.data_ptr(), shape, batch, precision, stream)
    torch.cuda.synchronize()
    return grad.cpu(), loss.cpu()


CASES = {
    "fused 2D": [((256, 256), 8, torch.complex64), ((64, 64), 2, torch.complex64)],
    "fused 3D": [((32, 32, 32), 3, torch.complex64), ((16, 16, 16), 1, torch.complex64)],
    "generic fp32": [((48, 40), 4, torch.complex64), ((24, 20), 2, torch.complex64)],
    "fp64": [((32, 32), 3, torch.complex128), ((16, 24), 2, torch.complex128)],
}


@pytest.mark.parametrize("name", list(CASES))
def test_noalloc_equals_the_allocating_entry(name):
    for k, (shape, batch, dtype) in enumerate(CASES[name]):
        precision = 1 if dtype == torch.complex128 else 0
        x, amp = _inputs(shape, batch, dtype, seed=k)
        want_grad, want_loss = _call(_qpch.amplitude_loss, x, amp, shape, batch, precision)
        got_grad, got_loss = _call(_qpch.amplitude_loss_noalloc, x, amp, shape, batch, precision)
        assert torch.isfinite(want_loss).all() and torch.isfinite(torch.view_as_real(want_grad)).all()
        assert torch.equal(got_loss, want_loss), (name, shape, batch)
        assert torch.equal(got_grad, want_grad), (name, shape, batch)


def test_noalloc_refuses_more_than_its_scratch():
    shape, batch = (512, 512), 16  # 4.2 million complex64 elements: 32 MiB of wave plus the partials
    x, amp = _inputs(shape, batch, torch.complex64, seed=3)
    grad = torch.zeros_like(x)
    loss = torch.zeros(batch, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="bytes of scratch"):
        _qpch.amplitude_loss_noalloc(x.data_ptr(), amp.data_ptr(), 0, grad.data_ptr(), loss.data_ptr(), shape,
                                     batch, 0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert torch.count_nonzero(loss) == 0  # nothing was queued
    # the allocating entry point takes it
    want_grad, want_loss = _call(_qpch.amplitude_loss, x, amp, shape, batch, 0)
    assert torch.isfinite(want_loss).all()
