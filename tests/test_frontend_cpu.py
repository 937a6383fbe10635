"""Frontend behaviour that needs no GPU: construction, plugin loading,
validation messages (kymatio/scattering2d/frontend, kymatio/frontend,
tests/scattering2d/test_torch_scattering2d.py:224-235,
tests/scattering2d/test_frontend_scattering2d.py), the 1D setup and
meta()/output_size(), and the no-CPU-fallback rule."""
from collections import namedtuple

import numpy as np
import pytest
import torch

import kymatio_amd
from kymatio_amd.torch import Scattering1D, Scattering2D
from oracle import filters1d


def test_fake_backend_not_supported():
    fake = namedtuple('backend', ['name'])
    fake.name = 'fake'
    with pytest.raises(ImportError, match='not supported'):
        Scattering2D(2, shape=(10, 10), backend=fake)


def test_backend_string_rules():
    with pytest.raises(ImportError, match='can not be called'):
        Scattering2D(2, shape=(10, 10), backend='doesnotexist')
    with pytest.raises(ImportError, match='not found'):
        Scattering2D(2, shape=(10, 10), backend='torch_nothing')
    assert Scattering2D(2, shape=(10, 10), backend='torch_hip').backend.fused


def test_frontend_entry():
    with pytest.raises(RuntimeError, match='is not valid'):
        kymatio_amd.Scattering2D(2, shape=(10, 10), frontend='doesnotexist')
    assert isinstance(kymatio_amd.Scattering2D(2, shape=(10, 10), frontend='torch'), Scattering2D)
    assert isinstance(kymatio_amd.Scattering1D(3, 64, 2, frontend='torch'), Scattering1D)


def test_frontend_entry_defaults_to_numpy():
    """kymatio/frontend/entry.py:20-21: no frontend= argument means NumPy."""
    from kymatio_amd.numpy import HarmonicScattering3D as NumPy3D
    from kymatio_amd.numpy import Scattering1D as NumPy1D
    from kymatio_amd.numpy import Scattering2D as NumPy2D
    assert type(kymatio_amd.Scattering2D(2, shape=(10, 10))) is NumPy2D
    assert type(kymatio_amd.Scattering1D(3, 64, 2)) is NumPy1D
    assert type(kymatio_amd.HarmonicScattering3D(1, (8, 8, 8))) is NumPy3D


def test_smallest_dimension():
    with pytest.raises(RuntimeError, match='smallest dimension'):
        Scattering2D(10, shape=(10, 10))


def test_buffers_match_reference_layout():
    sc = Scattering2D(3, (32, 32))
    bufs = dict(sc.named_buffers())
    assert len(bufs) == 3 + 8 * (1 + 2 + 2)
    assert bufs['tensor0'].shape == (48, 48, 1) and bufs['tensor0'].dtype == torch.float32
    assert 'tensor0' in sc.state_dict()


def test_cpu_input_raises_no_fallback():
    sc = Scattering2D(2, (16, 16))
    with pytest.raises(TypeError, match='HIP device'):
        sc(torch.zeros(16, 16))
    s1 = Scattering1D(3, 64, 2)
    with pytest.raises(TypeError, match='HIP device'):
        s1(torch.zeros(64))


@pytest.mark.parametrize('J,N,Q,kw', [(6, 4096, 8, {}), (6, 512, 16, {}), (5, 1000, 4, {}), (6, 512, 16, {'T': 16}),
                                      (6, 512, 16, {'stride': 4}), (5, 512, (12, 2), {})])
def test_1d_setup_matches_oracle(J, N, Q, kw):
    sc = Scattering1D(J, N, Q, **kw)
    ref = filters1d.setup(J, N, Q, kw.get('T'), kw.get('stride'))
    assert sc._N_padded == ref['N_pad']
    assert (sc.pad_left, sc.pad_right) == (ref['pad_left'], ref['pad_right'])
    assert sc.ind_start == ref['ind_start'] and sc.ind_end == ref['ind_end']
    assert [p['j'] for p in sc.psi1_f] == [p['j'] for p in ref['psi1']]
    for a, b in zip(sc.psi1_f, ref['psi1']):
        for la, lb in zip(a['levels'], b['levels']):
            assert torch.equal(la, torch.from_numpy(lb).float().view(-1, 1))


def test_1d_meta_and_output_size():
    sc = Scattering1D(6, 4096, 8)
    assert sc.output_size() == 126
    assert sc.output_size(detail=True) == (1, 38, 87)
    meta = sc.meta()
    assert meta['n'].shape == (126, 2)
    assert np.isnan(meta['n'][0]).all()
    assert meta['key'][1] == (0,) and meta['key'][-1] == (32, 6)


def test_1d_validation():
    with pytest.raises(ValueError, match='exactly one element'):
        Scattering1D(6, (5, 6), 8)
    with pytest.raises(ValueError, match='1-tuple'):
        Scattering1D(6, 1.5, 8)
    with pytest.raises(ValueError, match='Q must always be >= 1'):
        Scattering1D(6, 512, 0).Q
    sc = Scattering1D(6, 512, 8, T=0, out_type='array')
    with pytest.raises(ValueError, match='Please set out_type to'):
        sc._check_runtime_args()
    with pytest.raises(ValueError, match='stride must be a power of two'):
        Scattering1D(6, 512, 8, stride=3).log2_stride
