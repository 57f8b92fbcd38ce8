"""The hand-written adjoint of the amplitude loss on the GPU: gradcheck against finite
differences in FP64, the gradient against torch autograd of the same loss, the loss
against the oracle, and a short Adam refinement."""

import numpy as np
import pytest

import oracle
import synthetic

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ag():
    from qoraphasecu_amd import autograd

    assert torch.cuda.is_available()
    return autograd


def torch_reference(x, amp_fft, dims):
    """sum over usable pixels (|F x| - A)^2 with torch ops, A < 0 = masked (CPU, FP64)."""
    xc = x.detach().cpu().to(torch.complex128).requires_grad_(True)
    a = amp_fft.detach().cpu().to(torch.float64)
    y = torch.fft.fftn(xc, dim=dims, norm="ortho")
    keep = (a >= 0).to(torch.float64)
    loss = (((y.abs() - a.clamp(min=0)) ** 2) * keep).sum(dim=dims)
    loss.sum().backward()
    return loss.detach(), xc.grad


def rand_objects(shape, batch, dtype, seed=0):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn((batch, *shape), generator=g, dtype=torch.float64) + 1j * torch.randn(
        (batch, *shape), generator=g, dtype=torch.float64)
    return x.to(dtype)


@pytest.mark.parametrize("shape", [(32, 32), (24, 20), (8, 4, 16)])
def test_gradcheck_fp64(ag, shape):
    rng = np.random.default_rng(1)
    amp = torch.tensor(rng.uniform(0.5, 3.0, shape), dtype=torch.float64)
    amp.view(-1)[3] = -1.0  # one masked pixel
    amp = amp.cuda()
    x = rand_objects(shape, 2, torch.complex128).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda z: ag.AmplitudeLoss.apply(z, amp), (x,), eps=1e-6, atol=1e-6,
                                    rtol=1e-5)


@pytest.mark.parametrize("shape,dtype,tol", [((256, 256), torch.complex64, 1e-4), ((64, 48), torch.complex64, 1e-4),
                                             ((32, 32, 32), torch.complex64, 1e-4), ((32, 32), torch.complex128, 1e-12)])
def test_gradient_matches_torch_autograd(ag, shape, dtype, tol):
    obj, sup, inten = synthetic.triangle(shape, shape[-1] // 4)
    amp = ag.fft_order_amplitude(inten, mask=np.arange(inten.size).reshape(shape) % 97 == 0,
                                 dtype=torch.float32 if dtype == torch.complex64 else torch.float64)
    x = rand_objects(shape, 16 if len(shape) == 2 and shape[0] == 256 else 3, dtype).cuda().requires_grad_(True)
    loss = ag.amplitude_loss(x, amp)
    loss.sum().backward()
    dims = tuple(range(1, x.dim()))
    ref_loss, ref_grad = torch_reference(x, amp, dims)
    got = x.grad.detach().cpu().to(torch.complex128)
    rel_grad = (torch.linalg.vector_norm(got - ref_grad) / torch.linalg.vector_norm(ref_grad)).item()
    rel_loss = ((loss.cpu() - ref_loss).abs() / ref_loss).max().item()
    print(f"adjoint {shape} x {x.shape[0]} {dtype}: gradient vs FP64 autograd {rel_grad:.2e} (norm-relative), "
          f"loss {rel_loss:.2e}")
    assert rel_grad <= tol, rel_grad
    assert rel_loss <= tol, rel_loss


def test_per_object_amplitude(ag):
    shape = (64, 64)
    amps = []
    for seed in range(3):
        _, _, inten = synthetic.triangle(shape, 12 + 4 * seed)
        amps.append(ag.fft_order_amplitude(inten, ndim=2))
    amp = torch.stack(amps)
    x = rand_objects(shape, 3, torch.complex64).cuda().requires_grad_(True)
    loss = ag.amplitude_loss(x, amp)
    loss.sum().backward()
    for k in range(3):
        ref_loss, ref_grad = torch_reference(x[k:k + 1], amp[k], (1, 2))
        assert abs(loss[k].item() - ref_loss.item()) <= 1e-4 * ref_loss.item()
        g = x.grad[k:k + 1].detach().cpu().to(torch.complex128)
        assert (torch.linalg.vector_norm(g - ref_grad) / torch.linalg.vector_norm(ref_grad)).item() < 1e-4


def test_loss_matches_oracle(ag):
    obj, sup, inten = synthetic.triangle((48, 40), 12)
    amp_c, mask, _ = oracle.make_problem(inten, sup)
    x = synthetic.random_field(obj.size, 5).reshape(obj.shape)
    _, num = oracle.modulus_num(oracle.fft(oracle.to_fft_order(x), True), oracle.to_fft_order(amp_c),
                                oracle.to_fft_order(mask))
    amp = ag.fft_order_amplitude(inten)
    loss = ag.amplitude_loss(torch.tensor(x[None]).cuda(), amp)
    assert abs(loss.item() - num) <= 1e-5 * num


def test_adam_refinement_lowers_the_loss(ag):
    """Batch 16 of 256^2 objects from 50 RAAR iterations, Adam on the projected iterate."""
    import qoraphasecu_amd as qpc
    from qoraphasecu_amd import workloads

    obj, inten, mask = workloads.problem_2d(256, seed=0)
    ens = qpc.reconstruct(inten, mask=mask.astype(bool), recipe=qpc.Recipe(qpc.RAAR(50)), starts=16, backend="hip")
    x0 = np.stack([r.object for r in ens])
    support = np.stack([r.support for r in ens])
    out, hist = ag.refine(x0, inten, support, mask=mask.astype(bool), steps=100, lr=1e-2)
    h = hist.cpu().numpy()
    assert h.shape == (100, 16)
    assert np.all(h[-1] < h[0])


def test_graph_refinement_equals_eager_steps(ag):
    """The refinement captured in a CUDA graph (the loss through the no-allocation entry
    point, its adjoint and Adam's update, replayed per step) against the same steps run
    eagerly with the same Adam settings: every step's loss and the final object equal, bit
    for bit.  Run twice, with other GPU work and synchronisations between the runs."""
    from qoraphasecu_amd import workloads

    obj, inten, mask = workloads.problem_2d(256, seed=1)
    rng = np.random.default_rng(4)
    x0 = (rng.standard_normal((4, 256, 256)) + 1j * rng.standard_normal((4, 256, 256))).astype(np.complex64) * 30
    support = np.zeros((256, 256), bool)
    support[80:176, 80:176] = True
    eager, eh = ag.refine(x0, inten, support, mask=mask.astype(bool), steps=25, graph=False, capturable=True, fused=True)
    for run in range(2):
        busy = torch.randn(1 << 20, device="cuda").cumsum(0)  # unrelated GPU work between the runs
        torch.cuda.synchronize()
        out, h = ag.refine(x0, inten, support, mask=mask.astype(bool), steps=25, graph=True)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(busy).all())
        assert torch.equal(h, eh), (run, (h - eh).abs().max().item())
        assert torch.equal(out, eager), run
    # and within float rounding of the default eager optimizer (host-side step counter)
    plain, ph = ag.refine(x0, inten, support, mask=mask.astype(bool), steps=25)
    assert torch.allclose(h, ph, rtol=1e-4, atol=0.0)
