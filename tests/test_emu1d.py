"""1D workgroup bodies run single-threaded on the CPU (libkymatio_emu.so)
against the reference fixture, the golden vectors and the oracle, in every
averaging mode."""
import numpy as np
import pytest
import torch

from kymatio_amd.torch import Scattering1D
from oracle import scat1d
from tests.helpers import RTOL_CHANNEL, Emu, channel_rel_err, emu_scat1d, golden


@pytest.fixture(scope='module')
def emu():
    return Emu()


def test_fixture_1d(emu):
    d = golden('test_data_1d.npz')
    S = Scattering1D(int(d['J']), 512, int(d['Q']))
    out, C = emu_scat1d(emu, S, d['x'])
    out = out.reshape(d['Sx'].shape)
    assert np.allclose(out, d['Sx'])


@pytest.mark.parametrize('tag', ['cfg5', 'T16', 'stride4', 'Q12', 'o1', 'N1000'])
def test_golden_1d(emu, tag):
    G = golden('golden_scat1d.npz')
    x = G[tag + '_x']
    kw = {k: int(G[tag + '_' + k]) for k in ('T', 'stride', 'max_order') if tag + '_' + k in G.files}
    Q = G[tag + '_Q']
    Q = tuple(int(v) for v in Q) if Q.ndim else int(Q)
    S = Scattering1D(int(G[tag + '_J']), x.shape[-1], Q, **kw)
    out, _ = emu_scat1d(emu, S, x)
    assert emu.last_spec_launches1d >= 1  # FP32 forward nodes on the compile-time FFT lengths
    R = G[tag + '_Sx']
    assert channel_rel_err(out.reshape(R.shape), R) <= RTOL_CHANNEL


@pytest.mark.parametrize('T', [0, 'global'])
def test_unaveraged_and_global(emu, T):
    x = np.random.default_rng(1).standard_normal((2, 512)).astype(np.float32)
    S = Scattering1D(6, 512, 8, T=T, out_type='list')
    out, C = emu_scat1d(emu, S, x)
    ref = scat1d.Scattering1DTorchCPU(6, 512, 8, T=T, out_type='list')(torch.from_numpy(x))
    assert C == len(ref)
    offs = np.concatenate([[0], np.cumsum([r['coef'].shape[-1] for r in ref])])
    for i, r in enumerate(ref):
        o = out[:, offs[i]:offs[i + 1]]
        rr = r['coef'].numpy().reshape(o.shape)
        assert np.abs(o - rr).max() <= RTOL_CHANNEL * max(np.abs(rr).max(), 1e-30)


def test_fp64(emu):
    x = np.random.default_rng(2).standard_normal((2, 1024))
    S = Scattering1D(5, 1024, 4)
    out, C = emu_scat1d(emu, S, x, np.float64)
    ref = scat1d.Scattering1DTorchCPU(5, 1024, 4, dtype=torch.float64)(torch.from_numpy(x)).numpy()
    assert np.abs(out.reshape(ref.shape) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize('kw', [{}, {'T': 0, 'out_type': 'list'}, {'T': 'global', 'out_type': 'list'},
                                {'max_order': 1}, {'T': 16}, {'stride': 4}],
                         ids=['local', 'T0', 'global', 'order1', 'T16', 'stride4'])
def test_backward_matches_autograd_through_oracle(emu, kw):
    """FP64: hand-written 1D adjoint vs torch autograd through the restated
    reference, in every averaging mode."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, 512))
    S = Scattering1D(6, 512, 8, **kw)
    ref_mod = scat1d.Scattering1DTorchCPU(6, 512, 8, dtype=torch.float64,
                                          **{k: v for k, v in kw.items() if k != 'out_type'},
                                          out_type='list' if 'out_type' in kw else 'array')
    xt = torch.from_numpy(x).requires_grad_()
    ref = ref_mod(xt)
    coefs = [r['coef'] for r in ref] if isinstance(ref, list) else [ref[:, c] for c in range(ref.shape[1])]
    flat = torch.cat([c.reshape(2, -1) for c in coefs], dim=1)
    w = torch.from_numpy(rng.standard_normal(tuple(flat.shape)))
    (flat * w).sum().backward()
    out, C, gx = emu_scat1d(emu, S, x, np.float64, dS=w.numpy(), nbwd=3)
    assert np.abs(out - flat.detach().numpy()).max() <= 1e-12 * np.abs(flat.detach().numpy()).max()
    g = xt.grad.numpy()
    assert np.abs(gx - g).max() <= 1e-11 * np.abs(g).max()
