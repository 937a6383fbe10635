"""HarmonicScattering3D over the HIP ops, after the reference's
tests/scattering3d/test_torch_scattering3d.py: the fixture
test_data_3d.npz (the reference's own vectors), the closed form for a
single Gaussian, the backend ops against their definitions, and the checks
and messages.  Checks run on the CPU; transforms on the GPU."""
import numpy as np
import pytest
import torch

from kymatio_amd.scattering3d.backend.torch_backend import backend
from kymatio_amd.scattering3d.filter_bank import gaussian_3d, generate_weighted_sum_of_gaussians, solid_harmonic_3d
from kymatio_amd.torch import HarmonicScattering3D
from tests.helpers import golden

DEV = torch.device('cuda')


def relative_difference(a, b):
    return np.sum(np.abs(a - b)) / max(np.sum(np.abs(a)), np.sum(np.abs(b)))


def test_fft3d_errors():
    """test_torch_scattering3d.py:55-80."""
    with pytest.raises(TypeError, match='real'):
        backend.rfft(torch.randn(1, 4, 4, 4, 2))
    with pytest.raises(TypeError, match='complex'):
        backend.ifft(torch.randn(1, 4, 4, 4, 1))
    with pytest.raises(RuntimeError, match='must be contiguous'):
        backend.rfft(torch.randn(4, 4, 4, 1)[::2, ::2, ::2])
    with pytest.raises(RuntimeError, match='must be contiguous'):
        backend.ifft(torch.randn(4, 4, 4, 2)[::2, ::2, ::2])


def test_frontend_checks():
    S = HarmonicScattering3D(J=2, shape=(16, 16, 16))
    for k in range(3):
        with pytest.raises(RuntimeError, match='at least three'):
            S(torch.zeros((16, 16, 16)[:k]))
    with pytest.raises(RuntimeError, match='spatial size'):
        S(torch.zeros(15, 16, 16))
    with pytest.raises(TypeError, match='input should be'):
        S.scattering(np.zeros((16, 16, 16)))
    assert sorted(k for k, _ in S.named_buffers())[:3] == ['tensor0', 'tensor1', 'tensor2']


def test_filters_closed_forms():
    """l = 0 solid harmonics are the Gaussians; the order-l filters of every
    m share the radial profile: sum_m |psi_lm|^2 is isotropic, equal to
    (2l+1)/(4 pi) |r^l g|^2 norm^2 (addition theorem of spherical harmonics)."""
    g = gaussian_3d(16, 16, 16, 2.0)
    np.testing.assert_allclose(solid_harmonic_3d(16, 16, 16, 2.0, 0)[0].real, g, rtol=1e-6, atol=1e-7)
    l = 2
    f = solid_harmonic_3d(16, 16, 16, 2.0, l)
    p = (np.abs(f) ** 2).sum(0)
    grid = np.fft.ifftshift(np.mgrid[-8:8, -8:8, -8:8].astype('float32'), axes=(1, 2, 3)) * (2 * np.pi / 16)
    r2 = (grid ** 2).sum(0)
    # equal power at points of equal radius (permuted axes)
    np.testing.assert_allclose(p, p.transpose(1, 2, 0), rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(p, p.transpose(2, 0, 1), rtol=1e-4, atol=1e-9)
    assert p[r2 == 0].max() == 0


@pytest.mark.gpu
def test_backend_ops():
    """test_torch_scattering3d.py:41-50, 85-148 and the definitions of
    modulus_rotation / compute_integrals."""
    x = torch.zeros(1, 32, 32, 32, 1, device=DEV)
    x[0, 3, 5, 7, 0] = 2.0
    x[0, 0, 0, 0, 0] = 1.0
    y = backend.rfft(x)
    assert abs(y[0, 0, 0, 0, 0].item() - 3.0) < 1e-6
    xr = torch.randn(2, 8, 16, 4, 1, device=DEV)
    ref = torch.view_as_real(torch.fft.fftn(xr[..., 0].cpu().double(), dim=(-1, -2, -3)))
    assert (backend.rfft(xr).cpu().double() - ref).abs().max() < 1e-4 * ref.abs().max()
    z = torch.randn(2, 8, 16, 4, 2, device=DEV)
    back = backend.ifft(backend.rfft(xr))
    assert (back[..., 0] - xr[..., 0]).abs().max() < 1e-5 and back[..., 1].abs().max() < 1e-5
    a = torch.zeros(2, 3, 4, 2, device=DEV)
    a[..., 0], a[..., 1] = 2, 3
    b = torch.zeros_like(a)
    b[..., 0], b[..., 1] = 4, 5
    prod = backend.cdgmm3d(a, b)
    assert (prod[..., 0] - (8 - 15)).abs().max() < 1e-7 and (prod[..., 1] - (10 + 12)).abs().max() < 1e-7
    m1 = backend.modulus_rotation(z)
    m2 = backend.modulus_rotation(z * 2, m1)
    assert (m1[..., 0] - z.norm(dim=-1)).abs().max() < 1e-5
    assert (m2[..., 0] - torch.sqrt(5 * (z ** 2).sum(-1))).abs().max() < 1e-4
    u = torch.rand(3, 8, 8, 8, 1, device=DEV)
    I = backend.compute_integrals(u, [0.5, 1., 2.])
    for q, p in enumerate([0.5, 1., 2.]):
        r = (u.double() ** p).reshape(3, -1).sum(1)
        assert ((I[:, q].double() - r).abs() / r).max() < 1e-6


@pytest.mark.gpu
def test_against_standard_computations():
    """test_torch_scattering3d.py:150-203 on the reference's fixture."""
    d = golden('test_data_3d.npz')
    x = torch.from_numpy(d['x']).to(DEV)
    ref = d['Sx']
    J, L, P = int(d['J']), int(d['L']), d['integral_powers']
    B, M = x.shape[0], x.shape[1]
    S = HarmonicScattering3D(J=J, shape=(M, M, M), L=L, sigma_0=1, method='integral', integral_powers=P,
                             max_order=2).to(DEV)
    order_0 = backend.compute_integrals(x, P).cpu().numpy().reshape(B, -1)
    orders_12 = S(x).cpu().numpy().reshape(B, -1)
    assert relative_difference(ref[:, :order_0.shape[1]], order_0) < 1e-6
    assert relative_difference(ref[:, order_0.shape[1]:], orders_12) < 1e-6
    assert orders_12.dtype == np.float32


@pytest.mark.gpu
def test_solid_harmonic_scattering_closed_form():
    """test_torch_scattering3d.py:206-237: one Gaussian of width 3 against
    wavelets of width 3 2^j gives S[j, l] = k^l, k = s_w / sqrt(s_w^2 + s_g^2)."""
    M, J, L = 128, 1, 3
    grid = np.fft.ifftshift(np.mgrid[-M // 2:M // 2, -M // 2:M // 2, -M // 2:M // 2].astype('float32'),
                            axes=(1, 2, 3))
    x = torch.from_numpy(generate_weighted_sum_of_gaussians(grid, np.zeros((1, 1, 3)), np.ones((1, 1)), 3.)).float()
    S = HarmonicScattering3D(J=J, shape=(M, M, M), L=L, sigma_0=3., max_order=1, method='integral',
                             integral_powers=[1]).to(DEV)
    s = S(x.to(DEV)).cpu()
    for j in range(J + 1):
        sw = 3. * 2 ** j
        k = sw / np.sqrt(sw ** 2 + 9.)
        for l in range(1, L + 1):
            assert abs(s[0, j, l, 0].item() - k ** l) / (1e-6 + abs(s[0, j, l, 0].item())) < 1e-4


@pytest.mark.gpu
def test_batch_shape_agnostic_and_rotation_flag():
    """test_torch_scattering3d.py:255-288; rotation_covariant=False uses m = 0 only."""
    S = HarmonicScattering3D(J=2, shape=(16, 16, 16)).to(DEV)
    Sx = S(torch.zeros(16, 16, 16, device=DEV))
    assert len(Sx.shape) == 3
    for shape in ((1, 16, 16, 16), (2, 16, 16, 16), (2, 2, 16, 16, 16), (2, 2, 2, 16, 16, 16)):
        out = S(torch.zeros(shape, device=DEV))
        assert out.shape == shape[:-3] + Sx.shape[-3:]
    x = torch.rand(2, 16, 16, 16, device=DEV)
    a = HarmonicScattering3D(J=2, shape=(16, 16, 16), rotation_covariant=False).to(DEV)(x)
    b = HarmonicScattering3D(J=2, shape=(16, 16, 16)).to(DEV)(x)
    assert a.shape == b.shape and torch.allclose(a[:, :, 0], b[:, :, 0])  # l = 0 has the single m = 0


@pytest.mark.gpu
def test_numpy_and_sklearn_3d():
    from kymatio_amd.numpy import HarmonicScattering3D as NumPy3D
    from kymatio_amd.sklearn import HarmonicScattering3D as Sk3D
    d = golden('test_data_3d.npz')
    x = d['x']
    S = NumPy3D(J=int(d['J']), shape=x.shape[1:], L=int(d['L']), integral_powers=d['integral_powers'])
    Sx = S(x)
    assert isinstance(Sx, np.ndarray)
    ref = d['Sx'][:, len(d['integral_powers']):]
    assert relative_difference(ref, Sx.reshape(x.shape[0], -1)) < 1e-6
    t = Sk3D(J=int(d['J']), shape=x.shape[1:], L=int(d['L']),
             integral_powers=d['integral_powers']).transform(x.reshape(x.shape[0], -1))
    assert np.allclose(t, Sx.reshape(x.shape[0], -1))
