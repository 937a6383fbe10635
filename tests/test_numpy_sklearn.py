"""NumPy and scikit-learn frontends over the HIP path, after the reference's
tests/scattering2d/test_numpy_scattering2d.py:14-121,
tests/scattering2d/test_sklearn_2d.py:8-28 and
tests/scattering1d/test_numpy_scattering1d.py (fixture vectors
test_data_2d.npz / test_data_1d.npz).  Checks and messages run on the CPU;
the transforms run on the GPU (no CPU fallback)."""
from collections import namedtuple

import numpy as np
import pytest

import kymatio_amd
from kymatio_amd.numpy import Scattering1D as NumPy1D
from kymatio_amd.numpy import Scattering2D as NumPy2D
from kymatio_amd.sklearn import Scattering1D as Transformer1D
from kymatio_amd.sklearn import Scattering2D as Transformer2D
from tests.helpers import golden


def test_entry_dispatch_and_params():
    S = kymatio_amd.Scattering2D(2, (16, 16), frontend='numpy')
    assert type(S) is NumPy2D and S.frontend_name == 'numpy'
    T = kymatio_amd.Scattering2D(2, (16, 16), frontend='sklearn')
    assert type(T) is Transformer2D
    p = T.get_params()
    assert p['J'] == 2 and tuple(p['shape']) == (16, 16) and p['L'] == 8 and p['max_order'] == 2
    assert T.fit(np.zeros((3, 256))) is T
    S1 = kymatio_amd.Scattering1D(4, 256, 4, frontend='numpy')
    assert type(S1) is NumPy1D
    assert S1.output_size() == S1._spec.output_size()
    with pytest.raises(RuntimeError, match='frontend'):
        kymatio_amd.Scattering2D(2, (16, 16), frontend='nope')


def test_numpy2d_errors():
    """test_numpy_scattering2d.py:84-121."""
    fake = namedtuple('backend', ['name'])
    fake.name = 'fake'
    with pytest.raises(ImportError, match='not supported'):
        NumPy2D(2, shape=(10, 10), backend=fake)
    with pytest.raises(RuntimeError, match='smallest dimension'):
        NumPy2D(10, shape=(10, 10))
    S = NumPy2D(3, (32, 32))
    with pytest.raises(TypeError, match='input should be'):
        S(None)
    with pytest.raises(TypeError, match='numpy array'):
        S([[0.0] * 32] * 32)
    with pytest.raises(RuntimeError, match='have at least two dimensions'):
        S(np.random.randn(32))
    with pytest.raises(RuntimeError, match='NumPy array must be of spatial size'):
        S(np.random.randn(31, 31))
    S = NumPy2D(3, (32, 32), pre_pad=True)
    with pytest.raises(RuntimeError, match='Padded array must be of spatial size'):
        S(np.random.randn(31, 31))
    S = NumPy2D(3, (32, 32), out_type='dict')
    with pytest.raises(RuntimeError, match="out_type must be one of 'array' or 'list'"):
        S(np.zeros((32, 32), np.float32))


@pytest.mark.gpu
def test_numpy2d_fixture_and_orders():
    """test_numpy_scattering2d.py:16-46 on the reference's fixture."""
    d = golden('test_data_2d.npz')
    x, S0, J = d['x'], d['Sx'], int(d['J'])
    S = NumPy2D(J, x.shape[2:], pre_pad=bool(d['pre_pad']))
    Sx = S(x)
    assert isinstance(Sx, np.ndarray) and Sx.dtype == np.float32
    assert np.allclose(Sx, S0)
    S1 = NumPy2D(J, x.shape[2:], pre_pad=bool(d['pre_pad']), max_order=1)(x)
    assert np.allclose(S1, S0[..., :S1.shape[-3], :, :])
    L = NumPy2D(J, x.shape[2:], out_type='list')(x)
    assert isinstance(L[0]['coef'], np.ndarray) and len(L) == S0.shape[-3]
    assert np.allclose(np.stack([c['coef'] for c in L], -3), S0)


@pytest.mark.gpu
def test_numpy2d_batch_shape_float64():
    """test_numpy_scattering2d.py:48-78 (float64 input runs in float64)."""
    S = NumPy2D(3, (32, 32))
    Sx = S(np.zeros((32, 32)))
    assert Sx.shape == (217, 4, 4) and Sx.dtype == np.float64
    for shape in ((1, 32, 32), (2, 32, 32), (2, 2, 32, 32), (2, 2, 2, 32, 32)):
        Sx = S(np.zeros(shape))
        assert Sx.shape == shape[:-2] + (217, 4, 4)
    x = np.random.default_rng(0).standard_normal((2, 32, 32))
    a = S(x)
    b = S(x.astype(np.float32))
    assert a.dtype == np.float64 and b.dtype == np.float32
    assert np.abs(a - b).max() <= 1e-5 * np.abs(a).max()


@pytest.mark.gpu
def test_sklearn_transformer_2d():
    """test_sklearn_2d.py:8-28."""
    d = golden('test_data_2d.npz')
    x, J = d['x'], int(d['J'])
    Sx = NumPy2D(J, x.shape[2:]).scattering(x)
    st = Transformer2D(J, x.shape[2:])
    t = st.transform(x.reshape(x.shape[0], -1))
    assert np.allclose(Sx.reshape(x.shape[0], -1), t)
    from sklearn.pipeline import Pipeline
    pipe = Pipeline([('scattering', Transformer2D(J, x.shape[2:]))])
    assert np.allclose(pipe.fit_transform(x.reshape(-1, 32 * 32)), Sx.reshape(-1, Sx[0, 0].size))


@pytest.mark.gpu
def test_numpy1d_fixture_and_sklearn():
    """tests/scattering1d/test_numpy_scattering1d.py:16-34 on test_data_1d.npz."""
    d = golden('test_data_1d.npz')
    x, S0 = d['x'], d['Sx']
    S = NumPy1D(int(d['J']), x.shape[-1], int(d['Q']))
    Sx = S(x)
    assert isinstance(Sx, np.ndarray) and Sx.shape == S0.shape
    assert np.allclose(Sx, S0)
    t = Transformer1D(int(d['J']), x.shape[-1], int(d['Q'])).transform(x)
    assert np.allclose(t, Sx.reshape(x.shape[0], -1))
    D = NumPy1D(int(d['J']), x.shape[-1], int(d['Q']), out_type='dict')(x)
    assert all(isinstance(v, np.ndarray) for v in D.values())
