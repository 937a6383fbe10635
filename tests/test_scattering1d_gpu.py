"""Scattering1D on the GPU (through the C ABI) against the reference's
fixture, golden vectors written from the reference, and the oracle.
Parity bar (FP32): per path, max |ours - ref| <= 1e-5 * max |ref| of that path."""
import math

import numpy as np
import pytest
import torch

import kymatio_amd
from kymatio_amd import workloads
from kymatio_amd.torch import Scattering1D
from oracle import scat1d
from tests.helpers import RTOL_CHANNEL, channel_rel_err, golden

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def test_sample_scattering():
    """tests/scattering1d/test_torch_scattering1d.py:80-112."""
    d = golden('test_data_1d.npz')
    x = torch.from_numpy(d['x']).to(DEV)
    S0 = torch.from_numpy(d['Sx']).to(DEV)
    sc = kymatio_amd.Scattering1D(int(d['J']), x.shape[-1], int(d['Q']), frontend='torch').to(DEV)
    Sx = sc(x)
    assert Sx.shape == S0.shape
    assert torch.allclose(Sx, S0)


@pytest.mark.parametrize('tag', ['cfg5', 'T16', 'stride4', 'Q12', 'o1', 'N1000'])
def test_golden(tag):
    G = golden('golden_scat1d.npz')
    x = G[tag + '_x']
    kw = {k: int(G[tag + '_' + k]) for k in ('T', 'stride', 'max_order') if tag + '_' + k in G.files}
    Q = G[tag + '_Q']
    Q = tuple(int(v) for v in Q) if Q.ndim else int(Q)
    print('case', tag, flush=True)
    sc = Scattering1D(int(G[tag + '_J']), x.shape[-1], Q, **kw).to(DEV)
    out = sc(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert out.shape == G[tag + '_Sx'].shape
    assert channel_rel_err(out, G[tag + '_Sx']) <= RTOL_CHANNEL


def test_workload_batch_properties():
    """Config size (256 signals of 4096): batch equals single-signal runs bit
    for bit, deterministic, matches the oracle on a subset."""
    x = torch.from_numpy(workloads.chirp_signals(256, 4096, seed=5)).to(DEV)
    sc = Scattering1D(6, 4096, 8).to(DEV)
    a = sc(x)
    assert a.shape == (256, 126, 64)
    assert torch.equal(a, sc(x))
    for i in (0, 255):
        assert torch.equal(sc(x[i:i + 1].contiguous())[0], a[i])
    # chunks rotate over 1-4 streams (the plan's own streams joined back to
    # the caller's): the same bits, also with many chunks per stream
    for streams, per_chunk in ((1, None), (3, None), (4, 40), (2, 40)):
        sc.streams, sc.signals_per_chunk = streams, per_chunk
        sc._plans.clear()
        assert torch.equal(sc(x * 1.0), a), (streams, per_chunk)
    sc.streams, sc.signals_per_chunk = 0, None
    sc._plans.clear()
    ref = scat1d.Scattering1DTorchCPU(6, 4096, 8)(x[:4].cpu())
    assert channel_rel_err(a[:4].cpu().numpy(), ref.numpy()) <= RTOL_CHANNEL


@pytest.mark.parametrize('T', [0, 'global'])
def test_unaveraged_and_global(T):
    x = np.random.default_rng(1).standard_normal((2, 512)).astype(np.float32)
    sc = Scattering1D(6, 512, 8, T=T, out_type='list').to(DEV)
    out = sc(torch.from_numpy(x).to(DEV))
    ref = scat1d.Scattering1DTorchCPU(6, 512, 8, T=T, out_type='list')(torch.from_numpy(x))
    assert len(out) == len(ref)
    for o, r in zip(out, ref):
        assert o['n'] == r['n']
        oc, rc = o['coef'].cpu().numpy(), r['coef'].numpy().reshape(o['coef'].shape)
        assert np.abs(oc - rc).max() <= RTOL_CHANNEL * max(np.abs(rc).max(), 1e-30)


def test_fp64():
    x = np.random.default_rng(2).standard_normal((2, 1024))
    sc = Scattering1D(5, 1024, 4).double().to(DEV)
    out = sc(torch.from_numpy(x).to(DEV)).cpu().numpy()
    ref = scat1d.Scattering1DTorchCPU(5, 1024, 4, dtype=torch.float64)(torch.from_numpy(x)).numpy()
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def test_computation_Ux_and_dict():
    """tests/scattering1d/test_torch_scattering1d.py:115-158."""
    rng = np.random.RandomState(42)
    sc = Scattering1D(6, 2 ** 12, 8, T=0, max_order=1, out_type='dict').to(DEV)
    x = torch.from_numpy(rng.randn(1, 2 ** 12)).float().to(DEV)
    s = sc(x)
    for k in range(len(sc.psi1_f)):
        assert (k,) in s.keys()
    sc.max_order = 2
    s = sc(x)
    count = 1
    for k1, f1 in enumerate(sc.psi1_f):
        assert (k1,) in s.keys()
        count += 1
        for k2, f2 in enumerate(sc.psi2_f):
            if f2['j'] > f1['j']:
                assert (k1, k2) in s.keys()
                count += 1
    assert count == len(s)
    with pytest.raises(ValueError, match='Please set out_type to'):
        sc.out_type = 'array'
        sc(x)


def test_coordinates_and_output_size():
    """tests/scattering1d/test_torch_scattering1d.py:190-283."""
    torch.manual_seed(42)
    sc = Scattering1D(6, 2 ** 12, 8, max_order=2).to(DEV)
    x = torch.randn(2, 2 ** 12, device=DEV)
    for mo in (1, 2):
        sc.max_order = mo
        sc.out_type = 'dict'
        d = sc(x)
        sc.out_type = 'array'
        v = sc(x)
        meta = sc.meta()
        assert len(d) == v.shape[1] == sc.output_size()
        for cc in range(v.shape[1]):
            assert torch.equal(v[:, cc], d[meta['key'][cc]])


def test_batch_shape_agnostic():
    """tests/scattering1d/test_torch_scattering1d.py:332-381."""
    S = Scattering1D(3, (1024,), 8).to(DEV)
    with pytest.raises(ValueError, match='at least one axis'):
        S(torch.zeros((), device=DEV))
    Sx = S(torch.zeros(1024, device=DEV))
    assert Sx.dim() == 2 and Sx.shape[-1] == 1024 / 2 ** 3
    n = Sx.shape[-2]
    for shp in ((1, 1024), (2, 1024), (2, 2, 1024), (2, 2, 2, 1024)):
        S.out_type = 'array'
        Sx = S(torch.zeros(shp, device=DEV))
        assert Sx.shape == shp[:-1] + (n, 128)
        S.out_type = 'dict'
        assert len(S(torch.zeros(shp, device=DEV))) == n


def test_shape_errors():
    with pytest.raises(ValueError, match='exactly one element'):
        Scattering1D(6, (5, 6), 8)
    with pytest.raises(ValueError, match='1-tuple'):
        Scattering1D(6, 1.5, 8)
