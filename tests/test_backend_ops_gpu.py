"""Per-op HIP backend (kym_op_*) against the reference's backend unit tests
(tests/scattering2d/test_torch_backend_2d.py, tests/scattering1d/
test_torch_backend_1d.py, tests/general/test_torch_backend.py), run on GPU
tensors, and the per-op core loop (backend='torch_ops') against the fused
pipeline."""
import numpy as np
import pytest
import torch

from kymatio_amd.scattering1d.backend.torch_backend import backend as B1
from kymatio_amd.scattering2d.backend.torch_backend import backend as B2
from kymatio_amd.torch import Scattering2D
from tests.helpers import golden

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def test_pad_2d():
    pad = B2.Pad((2, 2, 2, 2), (4, 4))
    x = torch.randn(1, 4, 4, device=DEV)
    z = pad(x)
    assert z.shape == (1, 8, 8, 1)
    assert torch.allclose(z[0, 2, 2], x[0, 0, 0])
    assert torch.allclose(z[0, 1, 0], x[0, 1, 2])
    assert torch.allclose(z[0, 1, 1], x[0, 1, 1])
    assert torch.allclose(z[0, 1, 2], x[0, 1, 0])
    assert torch.allclose(z[0, 1, 3], x[0, 1, 1])
    # pad == size: numpy 'reflect'
    x = torch.randn(3, 4, 4, device=DEV)
    z = B2.Pad((4, 4, 4, 4), (4, 4))(x)[..., 0].cpu().numpy()
    assert np.array_equal(z, np.pad(x.cpu().numpy(), ((0, 0), (4, 4), (4, 4)), mode='reflect'))


def test_unpad_2d():
    x = torch.randn(4, 4, 1, device=DEV)
    y = B2.unpad(x)
    assert y.shape == (2, 2)
    assert torch.equal(y[0, 0], x[1, 1, 0])
    assert torch.equal(y[0, 1], x[1, 2, 0])


def test_modulus():
    x = torch.rand(100, 10, 4, 2, device=DEV)
    y = B2.modulus(x)[..., 0]
    assert torch.allclose(y, torch.sqrt(torch.sum(x * x, 3)))
    with pytest.raises(TypeError, match='should be complex'):
        B2.modulus(x[..., 0].contiguous())
    with pytest.raises(RuntimeError, match='contiguous'):
        B2.modulus(x[::2, ::2])


def test_subsample_fourier_2d():
    x = torch.rand(100, 1, 128, 128, 2, device=DEV)
    y = torch.zeros(100, 1, 8, 8, 2, device=DEV)
    for m in range(16):
        for n in range(16):
            y += x[..., m * 8:(m + 1) * 8, n * 8:(n + 1) * 8, :]
    assert torch.allclose(y / 256, B2.subsample_fourier(x, k=16))
    with pytest.raises(TypeError, match='should be complex'):
        B2.subsample_fourier(x[..., 0].clone(), k=16)
    with pytest.raises(RuntimeError, match='must be contiguous'):
        B2.subsample_fourier(x[::2, ::2], k=16)


@pytest.mark.parametrize('real', [False, True])
def test_cdgmm(real):
    x = torch.rand(100, 128, 128, 2, device=DEV)
    f = torch.rand(128, 128, 2, device=DEV)
    if real:
        f[..., 1] = 0
    y = torch.stack([x[..., 0] * f[..., 0] - x[..., 1] * f[..., 1], x[..., 1] * f[..., 0] + x[..., 0] * f[..., 1]], -1)
    if real:
        f = f[..., :1].contiguous()
    assert torch.allclose(y, B2.cdgmm(x, f), atol=1e-7, rtol=1e-6)


def test_cdgmm_exceptions():
    with pytest.raises(RuntimeError, match='not compatible'):
        B2.cdgmm(torch.empty(3, 4, 5, 2, device=DEV), torch.empty(4, 3, 2, device=DEV))
    with pytest.raises(TypeError, match='should be complex'):
        B2.cdgmm(torch.empty(3, 4, 5, 1, device=DEV), torch.empty(4, 5, 1, device=DEV))
    with pytest.raises(TypeError, match='same dtype'):
        B2.cdgmm(torch.empty(3, 4, 5, 2, device=DEV), torch.empty(4, 5, 1, device=DEV).double())
    with pytest.raises(TypeError) as exc:
        B2.cdgmm(torch.empty(3, 4, 5, 2), torch.empty(4, 5, 1, device=DEV))
    assert 'input must be on gpu' in exc.value.args[0].lower()
    with pytest.raises(TypeError) as exc:
        B2.cdgmm(torch.empty(3, 4, 5, 2, device=DEV), torch.empty(4, 5, 1))
    assert 'input must be on cpu' in exc.value.args[0].lower()


@pytest.mark.parametrize('shape', [(4, 4), (34, 34), (40, 24), (272, 272), (160, 136)])
def test_fft_2d(shape):
    """Brute-force DFT at 4x4 (reference test) and numpy FP64 at the workload sizes."""
    rng = np.random.default_rng(0)
    x = rng.random((3,) + shape)
    xt = torch.from_numpy(x)[..., None].to(DEV)
    z = B2.rfft(xt)
    ref = np.fft.fft2(x)
    got = z[..., 0].cpu().numpy() + 1j * z[..., 1].cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert torch.allclose(B2.irfft(z), xt)
    assert torch.allclose(B2.ifft(z)[..., :1], xt)


def test_fft_2d_multipass():
    """The two-pass route (forced by a tiny LDS cap) equals the in-LDS one."""
    import ctypes
    from kymatio_amd import _lib
    x = torch.randn(2, 68, 72, 2, dtype=torch.float64, device=DEV)
    outs = []
    for cap in (0, 4096):
        for kind in (0, 1):
            y = torch.empty_like(x)
            _lib.check(_lib.lib().kym_op_fft(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), None, 2,
                                              68, 72, kind, 1, cap, _lib.stream_of(x.device)))
            outs.append(y)
    assert torch.allclose(outs[0], outs[2], rtol=1e-12, atol=1e-12)
    assert torch.allclose(outs[1], outs[3], rtol=1e-12, atol=1e-14)


def test_fft_exceptions():
    with pytest.raises(TypeError, match='real'):
        B2.rfft(torch.randn(4, 4, 2, device=DEV))
    with pytest.raises(TypeError, match='complex'):
        B2.ifft(torch.randn(4, 4, 1, device=DEV))
    with pytest.raises(TypeError, match='complex'):
        B2.irfft(torch.randn(4, 4, 1, device=DEV))
    with pytest.raises(RuntimeError, match='must be contiguous'):
        B2.rfft(torch.randn(4, 4, 1, device=DEV)[::2, ::2])


def test_pad_1d():
    torch.manual_seed(42)
    N = 128
    for pad_left in range(0, N - 16, 16):
        for pad_right in (pad_left, pad_left + 16):
            x = torch.randn(2, 4, N, device=DEV)
            xp = B1.pad(x, pad_left, pad_right)[..., 0]
            ref = torch.nn.functional.pad(x, (pad_left, pad_right), mode='reflect')
            assert torch.equal(xp, ref)
    with pytest.raises(ValueError, match='padding size'):
        B1.pad(x, x.shape[-1], 0)
    with pytest.raises(ValueError, match='padding size'):
        B1.pad(x, 0, x.shape[-1])


def test_subsample_fourier_1d():
    """ifft(periodize(fft x, 2^j)) == x[::2^j] (test_torch_backend_1d.py:139-172)."""
    rng = np.random.RandomState(42)
    J = 10
    x = rng.randn(2, 4, 2 ** J) + 1j * rng.randn(2, 4, 2 ** J)
    xf = torch.view_as_real(torch.from_numpy(np.fft.fft(x, axis=-1))).contiguous().to(DEV)
    for j in range(J + 1):
        sub = B1.subsample_fourier(xf, 2 ** j).cpu()
        xs = np.fft.ifft(torch.view_as_complex(sub.contiguous()).numpy(), axis=-1)
        assert np.allclose(x[:, :, ::2 ** j], xs)


def test_fft_1d():
    x_r = np.random.rand(4)
    I, K = np.meshgrid(np.arange(4), np.arange(4), indexing='ij')
    y_r = (x_r * np.exp(-2 * np.pi * 1j * K * I / 4)).sum(-1)
    xt = torch.from_numpy(x_r)[..., None].to(DEV)
    yt = torch.from_numpy(np.column_stack((y_r.real, y_r.imag))).to(DEV)
    z = B1.rfft(xt)
    assert torch.allclose(yt, z)
    assert torch.allclose(xt[..., 0], B1.ifft(z)[..., 0])
    assert torch.allclose(xt, B1.irfft(z))
    assert torch.allclose(z, B1.cfft(B1.ifft(z)))


def test_per_op_core_loop_matches_fused():
    d = golden('test_data_2d.npz')
    x = torch.from_numpy(d['x']).to(DEV)
    ops = Scattering2D(4, (32, 32), backend='torch_ops').to(DEV)
    Sx = ops(x)
    assert torch.allclose(Sx, torch.from_numpy(d['Sx']).to(DEV))
    fused = Scattering2D(4, (32, 32)).to(DEV)
    assert torch.allclose(Sx, fused(x), rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize('B,T,N,k', [(3, 64, 88, 4), (2, 128, 88, 1), (4, 16, 22, 1)])
@pytest.mark.parametrize('mode', ['modulus', 'complex'])
def test_filter_ifft_swap_equals_op_sequence(B, T, N, k, mode):
    """filter_ifft_swap (kym_op_filter_ifft_t) equals the ops it fuses --
    filter_ifft, swap_time_frequency and modulus -- to FP32 rounding, and
    numpy FP64 of the reference's sequence."""
    rng = np.random.default_rng(B * T + N + k)
    x = rng.standard_normal((B, T, N)) + 1j * rng.standard_normal((B, T, N))
    f = rng.standard_normal(N)
    xt = torch.view_as_real(torch.from_numpy(x.astype(np.complex64))).contiguous().to(DEV)
    ft = torch.from_numpy(f.astype(np.float32))[:, None].to(DEV)
    fused = B1.filter_ifft_swap(xt, ft, k, mode)
    seq = B1.swap_time_frequency(B1.filter_ifft(xt, ft, k, 'complex'))
    if mode == 'modulus':
        seq = B1.modulus(seq)
    assert fused.shape == seq.shape
    scale = seq.abs().max().item()
    assert (fused - seq).abs().max().item() <= 1e-6 * scale
    ref = np.fft.ifft((x * f).reshape(B, T, k, N // k).mean(axis=2), axis=2).transpose(0, 2, 1)
    got = fused.cpu().numpy().astype(np.float64)
    if mode == 'modulus':
        assert np.abs(got[..., 0] - np.abs(ref)).max() <= 1e-5 * np.abs(ref).max()
    else:
        assert np.abs(got[..., 0] + 1j * got[..., 1] - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize('shape,axis,k', [((3, 22, 4096, 1), -2, 32), ((2, 88, 2048, 1), -2, 16),
                                          ((5, 88, 128, 1), -3, 4), ((3, 22, 128, 1), -3, 1), ((2, 7, 256, 1), -2, 2)])
def test_lowpass_equals_op_sequence(shape, axis, k):
    """lowpass (kym_op_lowpass) equals rfft then filter_ifft's real mode
    along time, and swap, rfft, filter_ifft and swap along frequency -- the
    reference's time_averaging and frequency_averaging -- to FP32 rounding,
    and numpy FP64."""
    rng = np.random.default_rng(sum(shape) + k)
    x = rng.standard_normal(shape)
    N = shape[axis]
    f = rng.standard_normal(N)
    xt = torch.from_numpy(x.astype(np.float32)).to(DEV)
    ft = torch.from_numpy(f.astype(np.float32))[:, None].to(DEV)
    fused = B1.lowpass(xt, ft, k, axis)
    if axis == -2:
        seq = B1.filter_ifft(B1.rfft(xt), ft, k, 'real')
    else:
        seq = B1.swap_time_frequency(B1.filter_ifft(B1.rfft(B1.swap_time_frequency(xt)), ft, k, 'real'))
    assert fused.shape == seq.shape
    scale = seq.abs().max().item()
    assert (fused - seq).abs().max().item() <= 1e-6 * scale
    ax = x.ndim + axis
    X = np.fft.fft(x, axis=ax) * f.reshape([-1 if i == ax else 1 for i in range(x.ndim)])
    Xs = np.moveaxis(X, ax, -1)
    ref = np.fft.ifft(Xs.reshape(Xs.shape[:-1] + (k, N // k)).mean(axis=-2), axis=-1).real
    ref = np.moveaxis(ref, -1, ax)
    assert np.abs(fused.cpu().numpy() - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize('c', [1, 2])
def test_freq_scatter_equals_op_sequence(c):
    """freq_scatter (kym_op_freq_scatter) equals the ops it fuses -- swap,
    pad_frequency, cfft / rfft, then per wavelet filter_ifft, swap and
    modulus -- to FP32 rounding."""
    B, A, T, N, ks = 3, 10, 64, 88, [1, 2, 4, 4, 1]
    rng = np.random.default_rng(c)
    x = torch.from_numpy(rng.standard_normal((B, 1, A, T, c)).astype(np.float32)).to(DEV)
    fs = [torch.from_numpy(rng.standard_normal((N, 1)).astype(np.float32)).to(DEV) for _ in ks]
    fused = B1.freq_scatter(x, fs, ks, N)
    X_T = B1.pad_frequency(B1.swap_time_frequency(x), N - A)
    X_hat = B1.cfft(X_T) if c == 2 else B1.rfft(X_T)
    for y, f, k in zip(fused, fs, ks):
        seq = B1.modulus(B1.swap_time_frequency(B1.filter_ifft(X_hat, f, k, 'complex')))
        assert y.shape == seq.shape
        assert (y - seq).abs().max().item() <= 1e-6 * seq.abs().max().item()
