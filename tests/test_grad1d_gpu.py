"""Scattering1D gradients on the GPU: gradcheck, torch autograd through the
oracle in every averaging mode, and the reference test that the gradient is
non-zero (tests/scattering1d/test_torch_scattering1d.py:286-311)."""
import numpy as np
import pytest
import torch
from torch.autograd import gradcheck

from kymatio_amd.torch import Scattering1D
from oracle import scat1d

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def test_differentiability():
    torch.manual_seed(42)
    sc = Scattering1D(6, 2 ** 12, 8).to(DEV)
    x = torch.randn(2, 2 ** 12, requires_grad=True, device=DEV)
    torch.sum(torch.abs(sc(x))).backward()
    assert torch.max(torch.abs(x.grad)) > 0.


def test_gradcheck_small():
    sc = Scattering1D(3, 64, 2).double().to(DEV)
    x = torch.randn(1, 64, dtype=torch.float64, device=DEV, requires_grad=True)
    assert gradcheck(sc, x, nondet_tol=1e-9)


@pytest.mark.parametrize('kw', [{}, {'T': 0, 'out_type': 'list'}, {'T': 'global', 'out_type': 'list'},
                                {'max_order': 1}, {'stride': 4}], ids=['local', 'T0', 'global', 'order1', 'stride4'])
def test_matches_oracle_autograd(kw):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 512))
    sc = Scattering1D(6, 512, 8, **kw).double().to(DEV)
    xt = torch.from_numpy(x).to(DEV).requires_grad_()
    out = sc(xt)
    coefs = [o['coef'] for o in out] if isinstance(out, list) else [out[:, c] for c in range(out.shape[1])]
    flat = torch.cat([c.reshape(2, -1) for c in coefs], dim=1)
    w = torch.from_numpy(rng.standard_normal(tuple(flat.shape))).to(DEV)
    (flat * w).sum().backward()
    ref_mod = scat1d.Scattering1DTorchCPU(6, 512, 8, dtype=torch.float64,
                                          **{k: v for k, v in kw.items() if k != 'out_type'},
                                          out_type='list' if 'out_type' in kw else 'array')
    xr = torch.from_numpy(x).requires_grad_()
    ref = ref_mod(xr)
    rc = [r['coef'] for r in ref] if isinstance(ref, list) else [ref[:, c] for c in range(ref.shape[1])]
    (torch.cat([c.reshape(2, -1) for c in rc], dim=1) * w.cpu()).sum().backward()
    g, gr = xt.grad.cpu(), xr.grad
    assert (g - gr).abs().max() <= 1e-11 * gr.abs().max()


def test_workload_grad_fp32_vs_fp64():
    """J=6 Q=8 N=4096: FP32 gradient against the FP64 one (same kernels)."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 4096))
    out = []
    for dt in (torch.float32, torch.float64):
        sc = Scattering1D(6, 4096, 8).to(dt).to(DEV)
        xt = torch.from_numpy(x).to(dt).to(DEV).requires_grad_()
        (sc(xt) ** 2).sum().backward()
        out.append(xt.grad.double().cpu())
    assert (out[0] - out[1]).abs().max() <= 1e-4 * out[1].abs().max()
