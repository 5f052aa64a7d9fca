"""GPU tests of the torch.autograd.Function wrapper (ambient_image_processor_amd.autograd)
around the C-ABI adjoint: torch device pointers go straight to the library."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _vol(shape, seed):
    from ambient_image_processor_amd import phantoms
    rng = phantoms.SplitMix64(seed)
    clean = phantoms._boxes(rng, shape, 6, max(shape[2] - 6, 1), (3, 8), (0.1, 1.0), 0.0)
    f = clean + rng.normal(int(np.prod(shape)), 0.05).reshape(shape)
    return f.astype(np.float32), clean.astype(np.float32)


@pytest.mark.parametrize("device", ["cuda", "cpu"])
def test_autograd_matches_c_abi_adjoint(ctx, device):
    from ambient_image_processor_amd import autograd, native
    f0, c0 = _vol((12, 16, 20), 1)
    f1, c1 = _vol((12, 16, 20), 2)
    f = torch.tensor(np.stack([f0, f1]), device=device, requires_grad=True)
    clean = torch.tensor(np.stack([c0, c1]), device=device)
    a0 = torch.tensor(0.05, device=device, requires_grad=True)
    a1 = torch.tensor(0.02, device=device, requires_grad=True)
    lam = torch.tensor(1.0, device=device, requires_grad=True)
    u = autograd.tgv2_l2_denoise(f, a0, a1, lam, iterations=30, checkpoint_every=7)
    loss = torch.mean((u - clean) ** 2)
    loss.backward()
    # the same through the C ABI, volume by volume
    tau = float(1 / np.sqrt(20))
    p = native.default_params(lam=1.0, alpha0=0.05, alpha1=0.02, tau=tau, sigma=tau,
                              iterations=30, data_term=native.DATA_L2)
    n = f.numel()
    gsum = np.zeros(3)
    for b, (fb, cb) in enumerate(((f0, c0), (f1, c1))):
        ub = ctx.tgv2_denoise(fb, p)
        np.testing.assert_array_equal(u[b].detach().cpu().numpy(), ub)
        _, gf, gp = ctx.tgv2_denoise_adjoint(fb, p, (2.0 / n) * (ub - cb), 7)
        np.testing.assert_allclose(f.grad[b].cpu().numpy(), gf, rtol=1e-5, atol=1e-12)
        gsum += gp
    np.testing.assert_allclose([a0.grad.item(), a1.grad.item(), lam.grad.item()], gsum,
                               rtol=1e-5)


def test_autograd_directional_fd():
    """Central difference of the GPU forward along a unit-RMS direction
    (step 1e-3 per voxel) against the adjoint's directional derivative."""
    from ambient_image_processor_amd import autograd
    f0, c0 = _vol((16, 16, 16), 3)
    f = torch.tensor(f0, device="cuda", requires_grad=True)
    clean = torch.tensor(c0, device="cuda")
    u = autograd.tgv2_l2_denoise(f, 2.0, 1.0, 1.0, iterations=50, checkpoint_every=10)
    loss = torch.mean((u - clean) ** 2)
    loss.backward()
    d = torch.tensor(np.random.default_rng(0).standard_normal(f0.shape).astype(np.float32),
                     device="cuda")
    eps = 1e-3
    with torch.no_grad():
        lp = torch.mean((autograd.tgv2_l2_denoise(f + eps * d, 2.0, 1.0, 1.0, 50) - clean) ** 2)
        lm = torch.mean((autograd.tgv2_l2_denoise(f - eps * d, 2.0, 1.0, 1.0, 50) - clean) ** 2)
    fd = (lp.double() - lm.double()).item() / (2 * eps)
    ad = float(torch.sum(f.grad.double() * d.double()).item())
    print("directional adjoint", ad, "fd", fd, flush=True)
    assert abs(ad - fd) <= 1e-2 * abs(fd)
