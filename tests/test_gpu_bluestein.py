"""Axis lengths with a prime factor above 64, from 256 on, run Bluestein's chirp-z transform
inside the generic kernels (a zero-padded power-of-two convolution in LDS) instead of a
plain O(N^2) DFT per line.  The reference's CPU backend transforms any length
(src/cpu/fft.cpp:25-35, pocketfft, which uses the same algorithm for large primes).

Checked against the FP64 oracle: transforms (fp32 to 2e-6 relative, as every other length;
FP64 to 1e-12), one teacher-forced iteration of the fused generic passes (1e-5), the FP64
amplitude-loss gradient against torch, and the no-allocation entry point, which refuses a
Bluestein length until its tables exist and then gives the allocating entry's bits."""

import numpy as np
import pytest

import oracle
import recipes as R
import synthetic

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("shape", [(257, 6), (6, 257), (521, 4), (1021, 3), (2039, 2), (2, 2039), (1042, 2),
                                   (3, 257, 4), (4, 5, 389), (1021, 16)])
@pytest.mark.parametrize("forward", [True, False])
def test_bluestein_fft_matches_oracle(qpch, shape, forward):
    rng = np.random.default_rng(sum(shape))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    got = qpch.fft(x, forward)
    want = oracle.fft(x.astype(np.complex128), forward, "d")
    print(f"{shape} {'forward' if forward else 'inverse'}: {rel(got, want):.2e}")
    assert rel(got, want) < 2e-6, (shape, rel(got, want))


@pytest.mark.parametrize("shape", [(257, 16), (16, 521), (8, 4, 389)])
def test_bluestein_teacher_forced_iteration(qpch, shape):
    obj, sup, inten = synthetic.triangle(shape, 4)
    amp, mask, sup = oracle.make_problem(inten, sup, None)
    mask = mask.copy()
    mask.flat[3] = 1
    start = np.stack([synthetic.random_field(obj.size, 11).reshape(shape),
                      synthetic.random_field(obj.size, 12).reshape(shape)])
    f = oracle.to_fft_order
    for kind, ops in [("hio", [R.hio(1)]), ("raar", [R.raar(1)]), ("er", [R.er(1)])]:
        o = oracle.Solver(amp, mask, sup, starts=2, start=start)
        o.run(ops)
        ores = o.results()
        for schedule in ("graph", "unfused"):
            s = qpch.Solver(f(amp), f(mask), f(sup), starts=2, start=np.stack([f(x) for x in start]),
                            schedule=schedule)
            errs = R.run_abi(s, ops)
            xs = s.read_state()
            s.close()
            for k in range(2):
                assert abs(errs[0, k] - ores[k][2][0]) <= 1e-5 * ores[k][2][0], (kind, schedule, k)
                assert rel(xs[k], f(ores[k][0])) < 1e-5, (kind, schedule, k, rel(xs[k], f(ores[k][0])))


def test_bluestein_fp64_adjoint_matches_torch():
    import torch

    from qoraphasecu_amd import autograd

    shape = (257, 4)
    rng = np.random.default_rng(5)
    x = torch.tensor(rng.standard_normal((2, *shape)) + 1j * rng.standard_normal((2, *shape)), dtype=torch.complex128,
                     device="cuda", requires_grad=True)
    amp = torch.tensor(rng.uniform(0.5, 2.0, shape), dtype=torch.float64, device="cuda")
    loss = autograd.AmplitudeLoss.apply(x, amp).sum()
    loss.backward()
    xr = x.detach().cpu().requires_grad_(True)
    ref = ((torch.fft.fftn(xr, dim=(1, 2), norm="ortho").abs() - amp.cpu()) ** 2).sum()
    ref.backward()
    assert abs(float(loss.detach()) - float(ref.detach())) <= 1e-10 * float(ref.detach())
    assert float((x.grad.cpu() - xr.grad).abs().max() / xr.grad.abs().max()) < 1e-10


def test_noalloc_needs_the_tables_first(qpch):
    import torch

    shape, batch = (12, 1279), 2  # 1279 is prime: a length no other test uses
    g = torch.Generator().manual_seed(1)
    x = torch.complex(torch.randn((batch, *shape), generator=g), torch.randn((batch, *shape), generator=g)).cuda()
    amp = (torch.rand(shape, generator=g) * 2).cuda()

    def call(entry):
        grad = torch.empty_like(x)
        loss = torch.empty(batch, dtype=torch.float64, device="cuda")
        entry(x.data_ptr(), amp.data_ptr(), 0, grad.data_ptr(), loss.data_ptr(), shape, batch, 0,
              torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return grad.cpu(), loss.cpu()

    with pytest.raises(ValueError, match="Bluestein tables"):
        call(qpch.amplitude_loss_noalloc)
    want = call(qpch.amplitude_loss)
    got = call(qpch.amplitude_loss_noalloc)
    assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1])
