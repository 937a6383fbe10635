"""TimeFrequencyScattering (JTFS) against the reference's outputs on seeded
chirps (tests/golden/golden_jtfs.npz, written by
tests/golden/make_jtfs_golden.py from kymatio's torch frontend on the CPU),
and the constructor checks of
tests/scattering1d/test_timefrequency_scattering.py.  Filters and meta() run
on the CPU; the transform runs on the GPU over the HIP ops."""
import json

import numpy as np
import pytest
import torch

from kymatio_amd.torch import TimeFrequencyScattering
from tests.helpers import golden

G = golden('golden_jtfs.npz')
TAGS = ['time', 'joint', 'dictF0', 'listglobal', 'stride']


def _kw(tag):
    kw = json.loads(str(G[tag + '_kw']))
    if isinstance(kw.get('Q'), list):
        kw['Q'] = tuple(kw['Q'])
    return kw


@pytest.mark.parametrize('tag', TAGS)
def test_frequency_filters_and_meta(tag):
    S = TimeFrequencyScattering(**_kw(tag))
    phi, psis = S.filters_fr
    np.testing.assert_allclose(phi['levels'][0].numpy().reshape(-1), G[tag + '_phi_fr'], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(np.stack([p['levels'][0].numpy().reshape(-1) for p in psis]), G[tag + '_psi_fr'],
                               rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose([p['xi'] for p in psis], G[tag + '_psi_fr_xi'], rtol=1e-12)
    assert [list(k) for k in S.meta()['key']] == json.loads(str(G[tag + '_meta_keys']))


def test_constructor_checks():
    """test_timefrequency_scattering.py:60-108."""
    with pytest.raises(ValueError, match='Q_fr must be >= 1'):
        TimeFrequencyScattering(J=5, J_fr=2, Q=8, shape=512, Q_fr=0)
    with pytest.raises(NotImplementedError, match='Q_fr must be an integer or 1-tuple'):
        TimeFrequencyScattering(J=5, J_fr=2, Q=8, shape=512, Q_fr=(1, 2))
    S = TimeFrequencyScattering(J=5, J_fr=2, Q=8, shape=512, F=0, format='joint')
    with pytest.raises(ValueError, match="format='joint'"):
        S._check_runtime_args()
    S = TimeFrequencyScattering(J=5, J_fr=2, Q=8, shape=512, format='spam')
    with pytest.raises(ValueError, match="format must be 'time' or 'joint'"):
        S._check_runtime_args()
    S = TimeFrequencyScattering(J=5, J_fr=2, Q=8, shape=512, F=0, stride_fr=2)
    with pytest.raises(ValueError, match='incompatible'):
        S.log2_stride_fr
    with pytest.raises(TypeError, match='GPU'):
        TimeFrequencyScattering(J=5, J_fr=2, Q=8, shape=512)(torch.zeros(512))


def _close(ours, ref):
    ours = np.asarray(ours, np.float64)
    ref = np.asarray(ref, np.float64)
    assert ours.shape == ref.shape
    err = np.abs(ours - ref)
    tol = 1e-5 * np.abs(ref).max() + 1e-5 * np.abs(ref)
    assert (err <= tol).all(), float((err / (np.abs(ref).max() + 1e-30)).max())


@pytest.mark.gpu
@pytest.mark.parametrize('tag', TAGS)
def test_matches_reference(tag):
    S = TimeFrequencyScattering(**_kw(tag)).cuda()
    x = torch.from_numpy(G[tag + '_x']).cuda()
    with torch.no_grad():
        Sx = S(x)
    B = x.shape[0]
    if isinstance(Sx, dict):
        keys = list(Sx.keys())
        assert [list(k) for k in keys] == json.loads(str(G[tag + '_keys']))
        flat = np.concatenate([Sx[k].cpu().numpy().reshape(B, -1) for k in keys], axis=1)
        _close(flat, G[tag + '_flat'])
    elif isinstance(Sx, list):
        assert [list(p['n']) for p in Sx] == json.loads(str(G[tag + '_keys']))
        flat = np.concatenate([p['coef'].cpu().numpy().reshape(B, -1) for p in Sx], axis=1)
        _close(flat, G[tag + '_flat'])
    else:
        _close(Sx.cpu().numpy(), G[tag + '_Sx'])


@pytest.mark.gpu
def test_batch_shapes_and_float64():
    S = TimeFrequencyScattering(J=6, J_fr=2, Q=8, shape=1024).cuda()
    x = torch.from_numpy(G['time_x']).cuda()
    with torch.no_grad():
        a = S(x)
        b = S(x.reshape(2, 1, 1024))
        assert b.shape == (2, 1) + a.shape[1:]
        assert torch.equal(a, b.reshape(a.shape))
        d = S.double()(x.double())
    assert d.dtype == torch.float64
    assert (d.float() - a).abs().max() <= 1e-5 * a.abs().max()


@pytest.mark.gpu
def test_numpy_and_sklearn_frontends():
    """test_timefrequency_scattering.py:497-560 (numpy / sklearn)."""
    from kymatio_amd.numpy import TimeFrequencyScattering as NumPyJTFS
    from kymatio_amd.sklearn import TimeFrequencyScattering as SkJTFS
    x = G['time_x']
    Sx = NumPyJTFS(**_kw('time'))(x)
    assert isinstance(Sx, np.ndarray)
    _close(Sx, G['time_Sx'])
    t = SkJTFS(**_kw('time')).transform(x)
    assert np.allclose(t, Sx.reshape(x.shape[0], -1))


def _flat(Sx, B):
    if isinstance(Sx, dict):
        return torch.cat([v.reshape(B, -1) for v in Sx.values()], dim=1)
    if isinstance(Sx, list):
        return torch.cat([p['coef'].reshape(B, -1) for p in Sx], dim=1)
    return Sx.reshape(B, -1)


@pytest.mark.gpu
@pytest.mark.parametrize('tag', TAGS)
def test_graph_replays_equal_eager(tag):
    """The captured-graph route (default for inference) against the eager
    launches, bit for bit, over several replays with other GPU work and
    synchronisations in between, and with a second input."""
    S = TimeFrequencyScattering(**_kw(tag)).cuda()
    x = torch.from_numpy(G[tag + '_x']).cuda()
    B = x.shape[0]
    with torch.no_grad():
        S.graphs = False
        eager = _flat(S(x), B)
        eager2 = _flat(S(x * 0.5 + 1.0), B)
        S.graphs = True
        for rep in range(3):
            junk = torch.randn(1 << 20, device='cuda')
            (junk * 3).sum().item()
            torch.cuda.synchronize()
            got = _flat(S(x), B)
            assert torch.equal(got, eager), (tag, rep)
            got2 = _flat(S(x * 0.5 + 1.0), B)
            assert torch.equal(got2, eager2), (tag, rep)
    assert S._graph_cache.captures == 1
