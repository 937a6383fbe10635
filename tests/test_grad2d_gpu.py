"""Scattering2D gradients on the GPU: the hand-written backward against
torch.autograd.gradcheck (the reference's own test), FP64 central finite
differences on 16x16 patches, torch autograd through the oracle, and the
gradients the reference produced (golden, J=4 128x128)."""
import numpy as np
import pytest
import torch
from torch.autograd import gradcheck

from kymatio_amd.torch import Scattering2D
from oracle import scat2d
from tests.helpers import golden

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def test_gradcheck_reference_case():
    """tests/scattering2d/test_torch_scattering2d.py:237-248."""
    sc = Scattering2D(2, shape=(8, 8)).double().to(DEV)
    x = torch.rand(2, 1, 8, 8).double().to(DEV).requires_grad_()
    assert gradcheck(sc, x, nondet_tol=1e-5, fast_mode=True)


def test_gradcheck_full_small():
    sc = Scattering2D(2, shape=(12, 10), L=4).double().to(DEV)
    x = torch.randn(1, 12, 10, dtype=torch.float64, device=DEV, requires_grad=True)
    assert gradcheck(sc, x, nondet_tol=1e-9)


def _loss(sc, x, y):
    return ((sc(x) - sc(y)) ** 2).sum()


def test_finite_differences_fp64_patches():
    """J=4, 128x128 FP64: directional derivatives along random directions
    supported on 16x16 patches, central differences with h = 1e-6."""
    G = golden('golden_scat2d.npz')
    x = torch.from_numpy(G['grad_x'].astype(np.float64)).to(DEV)
    y = torch.from_numpy(G['grad_y'].astype(np.float64)).to(DEV)
    sc = Scattering2D(4, (128, 128)).double().to(DEV)
    xg = x.clone().requires_grad_()
    _loss(sc, xg, y).backward()
    g = xg.grad
    rng = np.random.default_rng(0)
    h = 1e-6
    with torch.no_grad():
        for trial in range(4):
            i0, j0 = rng.integers(0, 112, size=2)
            v = torch.zeros_like(x)
            v[..., i0:i0 + 16, j0:j0 + 16] = torch.from_numpy(rng.standard_normal((16, 16))).to(DEV)
            fd = (_loss(sc, x + h * v, y) - _loss(sc, x - h * v, y)) / (2 * h)
            an = (g * v).sum()
            print('patch (%d,%d): fd %.10e analytic %.10e' % (i0, j0, fd.item(), an.item()))
            assert abs(fd.item() - an.item()) <= 1e-6 * max(abs(an.item()), 1e-12) + 1e-9


@pytest.mark.parametrize('dtype,tol', [(np.float64, 1e-10), (np.float32, 1e-5)])
def test_matches_reference_autograd_golden(dtype, tol):
    """d/dx sum((S(x)-S(y))^2) at J=4 128x128, as torch autograd through the
    reference computed it (tests/golden/make_golden.py)."""
    G = golden('golden_scat2d.npz')
    tag = 'f64' if dtype == np.float64 else 'f32'
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    x = torch.from_numpy(G['grad_x'].astype(dtype)).to(DEV).requires_grad_()
    y = torch.from_numpy(G['grad_y'].astype(dtype)).to(DEV)
    sc = Scattering2D(4, (128, 128)).to(tdt).to(DEV)
    loss = _loss(sc, x, y)
    loss.backward()
    ref = G['grad_' + tag + '_dx']
    got = x.grad.cpu().numpy()
    rel = np.abs(got - ref).max() / np.abs(ref).max()
    print('loss %.10e ref %.10e, grad max rel err %.3e' % (loss.item(), float(G['grad_' + tag + '_loss']), rel))
    assert abs(loss.item() - float(G['grad_' + tag + '_loss'])) <= tol * abs(float(G['grad_' + tag + '_loss']))
    assert rel <= tol


def test_fp32_backward_matches_fp64_oracle_autograd():
    """FP32 hand-written backward at J=4 128x128 (the compile-time node and
    160x160 pass kernels, many threads per workgroup) against torch autograd
    through the FP64 oracle, for a loss linear in S with random weights
    (no cancellation between S(x) and S(y) to amplify rounding): max error
    <= 1e-5 of the gradient's peak, on two images."""
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2, 128, 128))
    ref_sc = scat2d.Scattering2DTorchCPU(4, (128, 128), dtype=torch.float64)
    xr = torch.from_numpy(x).requires_grad_()
    S = ref_sc(xr)
    w = rng.standard_normal(tuple(S.shape))
    (S * torch.from_numpy(w)).sum().backward()
    g_ref = xr.grad.numpy()
    sc = Scattering2D(4, (128, 128)).to(DEV)
    xt = torch.from_numpy(x.astype(np.float32)).to(DEV).requires_grad_()
    St = sc(xt)
    (St * torch.from_numpy(w.astype(np.float32)).to(DEV)).sum().backward()
    got = xt.grad.cpu().double().numpy()
    s_err = (St.detach().cpu().double().numpy() - S.detach().numpy())
    rel_s = np.abs(s_err).max() / np.abs(S.detach().numpy()).max()
    rel_g = np.abs(got - g_ref).max() / np.abs(g_ref).max()
    print('FP32 vs FP64 oracle at J=4 128x128: S max rel %.3e, dL/dx max rel %.3e' % (rel_s, rel_g))
    assert rel_s <= 1e-5
    assert rel_g <= 1e-5


def test_matches_oracle_autograd_multipass():
    """Multi-pass (big-node) backward route forced by a small LDS cap."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 40, 36))
    w = None
    sc = Scattering2D(3, (40, 36)).double().to(DEV)
    sc.lds_limit = 8192
    xt = torch.from_numpy(x).to(DEV).requires_grad_()
    S = sc(xt)
    w = torch.from_numpy(rng.standard_normal(tuple(S.shape))).to(DEV)
    (S * w).sum().backward()
    ref_sc = scat2d.Scattering2DTorchCPU(3, (40, 36), dtype=torch.float64)
    xr = torch.from_numpy(x).requires_grad_()
    (ref_sc(xr) * w.cpu()).sum().backward()
    assert torch.allclose(xt.grad.cpu(), xr.grad, rtol=1e-10, atol=1e-12 * xr.grad.abs().max().item())


def test_no_grad_path_unchanged():
    sc = Scattering2D(3, (32, 32)).to(DEV)
    x = torch.randn(3, 32, 32, device=DEV)
    a = sc(x)
    b = sc(x.clone().requires_grad_())
    assert torch.allclose(a, b.detach(), rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize('B', [3, 8])
def test_two_stream_training_matches_one_stream(B):
    """The training forward and the backward split the batch over two
    streams (each image's saved state at its own workspace offset): outputs
    and gradients equal those of the one-stream plan bit for bit, also for an
    odd batch (3 images: three one-image chunks alternating over the two
    streams; 8 images: two chunks of 4), on two backward passes of one
    forward."""
    rng = np.random.default_rng(21)
    x = torch.from_numpy(rng.standard_normal((B, 64, 64)).astype(np.float32)).to(DEV)
    res = []
    for streams in (1, 2):
        sc = Scattering2D(3, (64, 64)).to(DEV)
        sc.tune = [16, 0, 0, 0, 0, streams]
        xt = x.clone().requires_grad_()
        S = sc(xt)
        w = torch.from_numpy(rng.standard_normal(tuple(S.shape)).astype(np.float32)).to(DEV) if not res else res[0][2]
        grads = torch.autograd.grad((S * w).sum(), xt, retain_graph=True)[0], \
            torch.autograd.grad((S * w).sum(), xt)[0]
        res.append((S.detach(), grads, w))
    (S1, (g1a, g1b), _), (S2, (g2a, g2b), _) = res
    assert torch.equal(S1, S2)
    assert torch.equal(g1a, g2a) and torch.equal(g1b, g2b) and torch.equal(g2a, g2b)
