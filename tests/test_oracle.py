"""The oracle against the reference's own fixtures and the golden vectors
written from the reference (tests/golden/make_golden.py), plus closed-form
checks of the ops it is built from."""
import numpy as np
import pytest
import torch

from oracle import filters1d, filters2d, scat1d, scat2d
from tests.helpers import golden, golden_2d_cases


def test_fixture_2d_torch_and_numpy():
    """tests/scattering2d/test_data_2d.npz, checked as the reference checks it
    (tests/scattering2d/test_torch_scattering2d.py:45-77)."""
    d = golden('test_data_2d.npz')
    x = torch.from_numpy(d['x'])
    S = torch.from_numpy(d['Sx'])
    J = int(d['J'])
    sc = scat2d.Scattering2DTorchCPU(J, (32, 32), pre_pad=bool(d['pre_pad']))
    Sg = sc(x)
    assert torch.allclose(Sg, S)
    S1 = scat2d.Scattering2DTorchCPU(J, (32, 32), max_order=1)(x)
    assert torch.allclose(S1, S[..., :S1.shape[-3], :, :])
    Sn = scat2d.scattering2d_numpy(d['x'], J, filters=sc.filters_np)
    assert np.allclose(Sn, d['Sx'])


def test_fixture_1d():
    """tests/scattering1d/test_data_1d.npz (test_torch_scattering1d.py:80-112)."""
    d = golden('test_data_1d.npz')
    x = torch.from_numpy(d['x'])
    S = torch.from_numpy(d['Sx'])
    sc = scat1d.Scattering1DTorchCPU(int(d['J']), x.shape[-1], int(d['Q']))
    assert torch.allclose(sc(x), S)


@pytest.mark.parametrize('tag,P', [('P40', 40), ('P160', 160)])
def test_filters2d_golden(tag, P):
    g = golden('golden_filters2d.npz')
    J = int(g[tag + '_J'])
    fb = filters2d.filter_bank(P, P, J, 8)
    allf = [fb['phi']['levels'][r] for r in range(J)] + [l for p in fb['psi'] for l in p['levels']]
    sums = np.array([a.astype(np.float64).sum() for a in allf])
    assert np.array_equal(sums, g[tag + '_sum'])
    vals = np.stack([a.reshape(-1)[i] for a, i in zip(allf, g[tag + '_idx'])])
    assert np.array_equal(vals, g[tag + '_val'])


@pytest.mark.parametrize('case', golden_2d_cases()[:-1], ids=lambda c: c[0])
def test_scat2d_golden(case):
    tag, x, J, shape, kw, ref = case
    sc = scat2d.Scattering2DTorchCPU(J, shape, **kw)
    out = sc(torch.from_numpy(x)).numpy()
    assert out.shape == ref.shape
    assert np.allclose(out, ref, rtol=1e-5, atol=1e-7)


def test_periodize_closed_form():
    """Alias-sum closed form (tests/scattering2d/test_torch_backend_2d.py:106-167)."""
    x = torch.rand(3, 1, 64, 64, 2, dtype=torch.float64)
    y = torch.zeros(3, 1, 8, 8, 2, dtype=torch.float64)
    for m in range(8):
        for n in range(8):
            y += x[..., m * 8:(m + 1) * 8, n * 8:(n + 1) * 8, :]
    assert torch.allclose(scat2d.subsample_torch(x, 8), y / 64)


def test_periodize_subsamples_1d():
    """ifft(periodize(fft x, 2^j)) == x[::2^j] (tests/scattering1d/test_torch_backend_1d.py:139-172)."""
    rng = np.random.RandomState(0)
    x = rng.randn(2, 3, 256) + 1j * rng.randn(2, 3, 256)
    xf = torch.view_as_real(torch.from_numpy(np.fft.fft(x, axis=-1)))
    for j in range(9):
        sub = torch.view_as_complex(scat1d.subsample(xf, 2 ** j).contiguous()).numpy()
        assert np.allclose(np.fft.ifft(sub, axis=-1), x[..., ::2 ** j])


def test_pad_reflect_known_answer():
    """tests/scattering2d/test_torch_backend_2d.py:41-56."""
    x = torch.randn(1, 4, 4)
    z = scat2d.pad_torch(x, (2, 2, 2, 2), [4, 4])
    assert z.shape == (1, 8, 8, 1)
    assert torch.equal(z[0, 2, 2, 0], x[0, 0, 0])
    assert torch.equal(z[0, 1, 0, 0], x[0, 1, 2])
    assert torch.equal(z[0, 1, 1, 0], x[0, 1, 1])
    assert torch.equal(z[0, 1, 2, 0], x[0, 1, 0])
    # pad == size special case equals numpy 'reflect'
    x = torch.randn(2, 4, 4)
    z = scat2d.pad_torch(x, (4, 4, 4, 4), [4, 4])[..., 0].numpy()
    assert np.array_equal(z, np.pad(x.numpy(), ((0, 0), (4, 4), (4, 4)), mode='reflect'))


def test_modulus_gradient_closed_form():
    """tests/general/test_torch_backend.py:6-26 and the zero sub-gradient."""
    x = torch.randn(50, 4, 2, dtype=torch.float64, requires_grad=True)
    y = scat2d.modulus_torch(x)
    y.sum().backward()
    assert torch.allclose(x.grad, x.detach() / y.detach())
    z = torch.zeros(5, 2, dtype=torch.float64, requires_grad=True)
    scat2d.modulus_torch(z).sum().backward()
    assert torch.all(z.grad == 0)


def test_fft_brute_force_dft():
    """Brute-force DFT at the FFT lengths the workloads use (the reference
    pins only 4 x 4): torch.fft as used by the oracle agrees with the O(n^2) sum."""
    for n in (5, 10, 17, 34, 40):
        x = np.random.default_rng(n).standard_normal(n) + 0j
        k = np.arange(n)
        dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
        got = torch.fft.fft(torch.from_numpy(x)).numpy()
        assert np.allclose(got, dft, atol=1e-12)


def test_filters1d_golden():
    g = golden('golden_scat1d.npz')
    cfg = filters1d.setup(6, 4096, 8)
    banks = [('phi', [cfg['phi']]), ('psi1_', cfg['psi1']), ('psi2_', cfg['psi2'])]
    sums = []
    for pre, bank in banks:
        for p in bank:
            for lev in p['levels']:
                sums.append(np.asarray(lev, np.float32).astype(np.float64).sum())
    assert np.allclose(np.array(sums), g['f8192_sum'], rtol=1e-6, atol=1e-9)
    assert np.array_equal(np.array([p['j'] for p in cfg['psi1']]), g['f8192_j1'])
    assert np.array_equal(np.array([p['j'] for p in cfg['psi2']]), g['f8192_j2'])
    assert np.allclose(np.array([p['xi'] for p in cfg['psi1']]), g['f8192_xi1'], rtol=0, atol=0)


@pytest.mark.parametrize('tag', ['cfg5', 'T16', 'stride4', 'Q12', 'o1', 'N1000'])
def test_scat1d_golden(tag):
    G = golden('golden_scat1d.npz')
    x = G[tag + '_x']
    kw = {}
    for k in ('T', 'stride', 'max_order'):
        if tag + '_' + k in G.files:
            kw[k] = int(G[tag + '_' + k])
    Q = G[tag + '_Q']
    Q = tuple(int(v) for v in Q) if Q.ndim else int(Q)
    sc = scat1d.Scattering1DTorchCPU(int(G[tag + '_J']), x.shape[-1], Q, **kw)
    out = sc(torch.from_numpy(x)).numpy()
    assert np.allclose(out, G[tag + '_Sx'], rtol=1e-5, atol=1e-7)
