"""Per-op FFT bodies (kym_op_fft) run on the CPU harness, for every kind on
the in-LDS and the two-pass routes, against numpy FP64, with guard regions
after each buffer to catch out-of-bounds writes."""
import ctypes

import numpy as np
import pytest

from tests.helpers import Emu

GUARD = 4096


@pytest.fixture(scope='module')
def emu():
    e = Emu()
    f = e.lib.kym_emu_op_fft
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] + [ctypes.c_int] * 4 + [ctypes.c_longlong]
    f.restype = ctypes.c_int
    return e


def _guarded(n, dtype):
    buf = np.full(n + GUARD, 12345.0, dtype)
    return buf


def _run(emu, x_c, H, W, kind, cap, with_work=True):
    B = x_c.shape[0]
    n = B * H * W
    if kind == 2:
        xin = _guarded(n, np.float64)
        xin[:n] = x_c.real.reshape(-1)
    else:
        xin = _guarded(2 * n, np.float64)
        xin[:2 * n] = np.stack([x_c.real, x_c.imag], -1).reshape(-1)
    out_len = n if kind == 3 else 2 * n
    y = _guarded(out_len, np.float64)
    work = _guarded(2 * n, np.float64) if with_work else None
    rc = emu.lib.kym_emu_op_fft(xin.ctypes.data, y.ctypes.data, work.ctypes.data if work is not None else None, B,
                                H, W, kind, 1, cap)
    if rc:
        raise RuntimeError(emu.lib.kym_emu_last_error().decode())
    assert np.all(y[out_len:] == 12345.0), 'write past the output'
    assert np.all(xin[len(xin) - GUARD:] == 12345.0)
    if work is not None:
        assert np.all(work[2 * n:] == 12345.0), 'write past the work buffer'
    if kind == 3:
        return y[:n].reshape(B, H, W)
    v = y[:out_len].reshape(B, H, W, 2)
    return v[..., 0] + 1j * v[..., 1]


@pytest.mark.parametrize('H,W', [(4, 4), (34, 34), (40, 24), (68, 72), (17, 1), (272, 16)])
@pytest.mark.parametrize('cap', [160 * 1024, 8192], ids=['lds', 'twopass'])
@pytest.mark.parametrize('kind', [0, 1, 2, 3])
def test_op_fft(emu, H, W, cap, kind):
    rng = np.random.default_rng(H * 1000 + W)
    x = rng.standard_normal((2, H, W)) + (0 if kind == 2 else 1j) * rng.standard_normal((2, H, W))
    got = _run(emu, x, H, W, kind, cap)
    if kind in (0, 2):
        ref = np.fft.fft2(x)
    elif kind == 1:
        ref = np.fft.ifft2(x)
    else:
        ref = np.fft.ifft2(x).real
    assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1)


@pytest.mark.parametrize('n,B', [(88, 100), (1024, 9), (5, 3000), (136, 31)])
@pytest.mark.parametrize('kind', [0, 1, 2, 3])
def test_op_fft_1d_several_per_workgroup(emu, n, B, kind):
    """Short 1D transforms run several per workgroup (the last workgroup
    partly filled): every kind against numpy FP64."""
    rng = np.random.default_rng(n + B)
    x = rng.standard_normal((B, n, 1)) + (0 if kind == 2 else 1j) * rng.standard_normal((B, n, 1))
    got = _run(emu, x, n, 1, kind, 160 * 1024)
    if kind in (0, 2):
        ref = np.fft.fft(x, axis=1)
    elif kind == 1:
        ref = np.fft.ifft(x, axis=1)
    else:
        ref = np.fft.ifft(x, axis=1).real
    assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1)


def test_real_inverse_two_pass_needs_work(emu):
    x = np.ones((1, 68, 72)) + 0j
    with pytest.raises(RuntimeError, match='work buffer'):
        _run(emu, x, 68, 72, 3, 8192, with_work=False)


@pytest.mark.parametrize('N,k,B', [(88, 1, 100), (8192, 4, 3), (4096, 32, 5), (64, 2, 700), (136, 8, 9),
                                   (4096, 32, 64), (8192, 1, 2), (88, 4, 3000)])
@pytest.mark.parametrize('mode', [0, 1, 2])
def test_filter_ifft(emu, N, k, B, mode):
    """kym_op_filter_ifft's body on the CPU harness: ifft(P_k(x f)) against
    numpy FP64 of the reference's op sequence (cdgmm, subsample_fourier's
    mean over the k aliases, ifft), kept complex, real or as the modulus."""
    f = emu.lib.kym_emu_filter_ifft
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] * 2 + [ctypes.c_int] * 3
    f.restype = ctypes.c_int
    rng = np.random.default_rng(N + k + B + mode)
    x = rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))
    filt = rng.standard_normal(N)
    n = N // k
    xin = np.stack([x.real, x.imag], -1).reshape(-1).copy()
    out_len = B * n * (2 if mode == 0 else 1)
    y = _guarded(out_len, np.float64)
    assert f(xin.ctypes.data, filt.ctypes.data, y.ctypes.data, B, N, k, mode, 1) == 0
    assert np.all(y[out_len:] == 12345.0), 'write past the output'
    ref = np.fft.ifft((x * filt).reshape(B, k, n).mean(axis=1), axis=1)
    if mode == 0:
        got = y[:out_len].reshape(B, n, 2)
        got = got[..., 0] + 1j * got[..., 1]
    else:
        got = y[:out_len].reshape(B, n)
        ref = ref.real if mode == 1 else np.abs(ref)
    assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1)


@pytest.mark.parametrize('N,k,T,B', [(88, 4, 64, 3), (88, 1, 128, 2), (22, 1, 16, 5), (64, 2, 8, 7)])
@pytest.mark.parametrize('mode', [0, 2])
def test_filter_ifft_transposed(emu, N, k, T, B, mode):
    """kym_op_filter_ifft_t: the rows are (b, t) and the output comes out
    as (b, n, t) -- ifft(P_k(x f)) followed by the swap of the time and
    frequency axes, and the modulus for mode 2 -- against numpy FP64."""
    f = emu.lib.kym_emu_filter_ifft_t
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] * 2 + [ctypes.c_int] * 2 + [ctypes.c_longlong,
                                                                                          ctypes.c_int]
    f.restype = ctypes.c_int
    rng = np.random.default_rng(N + k + T + B + mode)
    x = rng.standard_normal((B, T, N)) + 1j * rng.standard_normal((B, T, N))
    filt = rng.standard_normal(N)
    n = N // k
    xin = np.stack([x.real, x.imag], -1).reshape(-1).copy()
    out_len = B * T * n * (2 if mode == 0 else 1)
    y = _guarded(out_len, np.float64)
    assert f(xin.ctypes.data, filt.ctypes.data, y.ctypes.data, B * T, N, k, mode, T, 1) == 0
    assert np.all(y[out_len:] == 12345.0), 'write past the output'
    ref = np.fft.ifft((x * filt).reshape(B, T, k, n).mean(axis=2), axis=2).transpose(0, 2, 1)
    if mode == 0:
        got = y[:out_len].reshape(B, n, T, 2)
        got = got[..., 0] + 1j * got[..., 1]
    else:
        got = y[:out_len].reshape(B, n, T)
        ref = np.abs(ref)
    assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1)
    assert f(xin.ctypes.data, filt.ctypes.data, y.ctypes.data, B * T + 1, N, k, mode, T, 1) != 0


@pytest.mark.parametrize('N,k,S,outer', [(4096, 32, 1, 3), (2048, 16, 1, 5), (256, 2, 1, 40), (88, 4, 128, 3),
                                         (22, 1, 128, 2), (88, 4, 5, 7), (64, 64, 1, 9)])
def test_lowpass(emu, N, k, S, outer):
    """kym_op_lowpass: irfft(P_k(rfft(x) f)) along the middle axis of
    (outer, N, S) real lines -- the reference's rfft, cdgmm,
    subsample_fourier and irfft -- against numpy FP64."""
    f = emu.lib.kym_emu_lowpass
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] * 3 + [ctypes.c_int] * 2
    f.restype = ctypes.c_int
    rng = np.random.default_rng(N + k + S + outer)
    x = rng.standard_normal((outer, N, S))
    filt = rng.standard_normal(N)
    n = N // k
    y = _guarded(outer * n * S, np.float64)
    assert f(x.ctypes.data, filt.ctypes.data, y.ctypes.data, outer, N, S, k, 1) == 0
    assert np.all(y[outer * n * S:] == 12345.0), 'write past the output'
    X = np.fft.fft(x, axis=1)
    ref = np.fft.ifft((X * filt[None, :, None]).reshape(outer, k, n, S).mean(axis=1), axis=1).real
    got = y[:outer * n * S].reshape(outer, n, S)
    assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1)


@pytest.mark.parametrize('N,k,B,nf,mode', [(256, 4, 3, 5, 2), (128, 1, 4, 3, 0), (64, 8, 2, 1, 1)])
def test_filter_ifft_bank(emu, N, k, B, nf, mode):
    """kym_op_filter_ifft_g with a bank of nf filters: row r = x row r / nf
    with filter r mod nf, against numpy FP64 of ifft(P_k(x f))."""
    f = emu.lib.kym_emu_filter_ifft_g
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] * 2 + [ctypes.c_int] * 7
    f.restype = ctypes.c_int
    rng = np.random.default_rng(N + k + B + nf)
    x = rng.standard_normal((B, N)) + 1j * rng.standard_normal((B, N))
    bank = rng.standard_normal((nf, N))
    n = N // k
    xin = np.stack([x.real, x.imag], -1).reshape(-1).copy()
    c = 2 if mode == 0 else 1
    out_len = B * nf * n * c
    y = _guarded(out_len, np.float64)
    assert f(xin.ctypes.data, bank.ctypes.data, y.ctypes.data, B * nf, N, k, mode, nf, 0, 0, 0, 1) == 0
    assert np.all(y[out_len:] == 12345.0), 'write past the output'
    ref = np.fft.ifft((x[:, None, :] * bank[None]).reshape(B, nf, k, n).mean(axis=2), axis=-1)
    if mode == 0:
        got = y[:out_len].reshape(B, nf, n, 2)
        got = got[..., 0] + 1j * got[..., 1]
    else:
        got = y[:out_len].reshape(B, nf, n)
        ref = ref.real if mode == 1 else np.abs(ref)
    assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1)


def test_filter_ifft_row_map(emu):
    """kym_op_filter_ifft_g's output row map: the rows of a (B, G) block land
    at [:, goff:goff + G] of a (B, gtot) output and nothing else is written."""
    f = emu.lib.kym_emu_filter_ifft_g
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] * 2 + [ctypes.c_int] * 7
    f.restype = ctypes.c_int
    B, G, gtot, goff, N, k = 3, 4, 9, 2, 64, 2
    n = N // k
    rng = np.random.default_rng(7)
    x = rng.standard_normal((B, G, N)) + 1j * rng.standard_normal((B, G, N))
    filt = rng.standard_normal(N)
    xin = np.stack([x.real, x.imag], -1).reshape(-1).copy()
    y = _guarded(B * gtot * n * 2, np.float64)
    y[:B * gtot * n * 2] = -7.0
    assert f(xin.ctypes.data, filt.ctypes.data, y.ctypes.data, B * G, N, k, 0, 1, G, gtot, goff, 1) == 0
    out = y[:B * gtot * n * 2].reshape(B, gtot, n, 2)
    ref = np.fft.ifft((x * filt).reshape(B, G, k, n).mean(axis=2), axis=-1)
    got = out[:, goff:goff + G, :, 0] + 1j * out[:, goff:goff + G, :, 1]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    untouched = np.ones(gtot, bool)
    untouched[goff:goff + G] = False
    assert np.all(out[:, untouched] == -7.0)
    assert np.all(y[B * gtot * n * 2:] == 12345.0)
    assert f(xin.ctypes.data, filt.ctypes.data, y.ctypes.data, B * G, N, k, 0, 1, G, G + 1, 2, 1) != 0


@pytest.mark.parametrize('N,k,outer', [(4096, 32, 3), (2048, 16, 5), (1024, 8, 6), (512, 4, 9), (256, 2, 20)])
def test_lowpass_compiled_lengths(emu, N, k, outer):
    """The compile-time low-pass (FP32, lines of 256 to 4096 down to 128
    outputs, several lines per workgroup with a partial last one) against
    numpy FP64 to FP32 rounding."""
    f = emu.lib.kym_emu_lowpass
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] * 3 + [ctypes.c_int] * 2
    f.restype = ctypes.c_int
    rng = np.random.default_rng(N + k + outer)
    x = rng.standard_normal((outer, N)).astype(np.float32)
    filt = rng.standard_normal(N).astype(np.float32)
    n = N // k
    y = _guarded(outer * n, np.float32)
    assert f(x.ctypes.data, filt.ctypes.data, y.ctypes.data, outer, N, 1, k, 0) == 0
    assert np.all(y[outer * n:] == 12345.0), 'write past the output'
    X = np.fft.fft(x.astype(np.float64), axis=1) * filt.astype(np.float64)
    ref = np.fft.ifft(X.reshape(outer, k, n).mean(axis=1), axis=1).real
    got = y[:outer * n].reshape(outer, n)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize('N,k,T,B,mode', [(88, 1, 64, 3, 2), (88, 4, 128, 2, 2), (88, 1, 32, 2, 0), (22, 1, 128, 3, 2),
                                          (88, 1, 0, 70, 1)])
def test_filter_ifft_compiled_lengths(emu, N, k, T, B, mode):
    """The compile-time filter_ifft (FP32, outputs of 88 or 22, transposed
    rows or natural ones with a partial last workgroup) against numpy FP64 to
    FP32 rounding."""
    f = emu.lib.kym_emu_filter_ifft_t
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] * 2 + [ctypes.c_int] * 2 + [ctypes.c_longlong,
                                                                                          ctypes.c_int]
    f.restype = ctypes.c_int
    rng = np.random.default_rng(N + k + T + B + mode)
    rows = B * T if T else B
    x = (rng.standard_normal((rows, N)) + 1j * rng.standard_normal((rows, N))).astype(np.complex64)
    filt = rng.standard_normal(N).astype(np.float32)
    n = N // k
    xin = np.stack([x.real, x.imag], -1).reshape(-1).copy()
    c = 2 if mode == 0 else 1
    y = _guarded(rows * n * c, np.float32)
    assert f(xin.ctypes.data, filt.ctypes.data, y.ctypes.data, rows, N, k, mode, T, 0) == 0
    assert np.all(y[rows * n * c:] == 12345.0), 'write past the output'
    ref = np.fft.ifft((x.astype(np.complex128) * filt).reshape(rows, k, n).mean(axis=1), axis=1)
    if T:
        ref = ref.reshape(B, T, n).transpose(0, 2, 1)
    if mode == 0:
        got = y[:rows * n * 2].reshape(ref.shape + (2,))
        got = got[..., 0] + 1j * got[..., 1]
    else:
        got = y[:rows * n].reshape(ref.shape)
        ref = ref.real if mode == 1 else np.abs(ref)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize('N,k,S,outer', [(88, 4, 128, 3), (22, 1, 128, 2), (88, 4, 64, 2)])
def test_lowpass_compiled_strided(emu, N, k, S, outer):
    """The compile-time low-pass along a strided axis (FP32, the padded
    frequency axis of 88 or 22 points) against numpy FP64 to FP32 rounding."""
    f = emu.lib.kym_emu_lowpass
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_longlong] * 3 + [ctypes.c_int] * 2
    f.restype = ctypes.c_int
    rng = np.random.default_rng(N + k + S + outer)
    x = rng.standard_normal((outer, N, S)).astype(np.float32)
    filt = rng.standard_normal(N).astype(np.float32)
    n = N // k
    y = _guarded(outer * n * S, np.float32)
    assert f(x.ctypes.data, filt.ctypes.data, y.ctypes.data, outer, N, S, k, 0) == 0
    assert np.all(y[outer * n * S:] == 12345.0), 'write past the output'
    X = np.fft.fft(x.astype(np.float64), axis=1) * filt.astype(np.float64)[None, :, None]
    ref = np.fft.ifft(X.reshape(outer, k, n, S).mean(axis=1), axis=1).real
    got = y[:outer * n * S].reshape(outer, n, S)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize('B,A,T,c,N,ks', [(2, 10, 64, 2, 88, [1, 2, 4, 4, 2, 1, 4]), (3, 38, 16, 1, 88, [4, 1, 2]),
                                          (2, 7, 40, 2, 24, [1, 3, 8])])
def test_freq_scatter(emu, B, A, T, c, N, ks):
    """kym_op_freq_scatter: x (B, A, T) zero-padded to N along A, transformed
    along A, then per wavelet |ifft(P_k(X f))| laid out (B, N / k, T) --
    against numpy FP64 of the reference's swap, pad, fft, cdgmm,
    subsample_fourier, ifft, swap and modulus (T not a multiple of the
    workgroup's t-block in the last case)."""
    f = emu.lib.kym_emu_freq_scatter
    f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_longlong] + [ctypes.c_int] * 5
    f.restype = ctypes.c_int
    rng = np.random.default_rng(B + A + T + c + N)
    x = rng.standard_normal((B, A, T)) + (1j * rng.standard_normal((B, A, T)) if c == 2 else 0)
    xin = (np.stack([x.real, x.imag], -1) if c == 2 else x.real).reshape(-1).copy()
    filts = [rng.standard_normal(N) for _ in ks]
    ys = [_guarded(B * (N // k) * T, np.float64) for k in ks]
    yp = (ctypes.c_void_p * len(ks))(*[y.ctypes.data for y in ys])
    fp = (ctypes.c_void_p * len(ks))(*[g.ctypes.data for g in filts])
    kp = (ctypes.c_int * len(ks))(*ks)
    assert f(xin.ctypes.data, yp, fp, kp, len(ks), B, A, T, c, N, 1) == 0
    xp = np.zeros((B, N, T), complex)
    xp[:, :A] = x
    X = np.fft.fft(xp, axis=1)
    for y, g, k in zip(ys, filts, ks):
        n = N // k
        assert np.all(y[B * n * T:] == 12345.0), 'write past the output'
        ref = np.abs(np.fft.ifft((X * g[None, :, None]).reshape(B, k, n, T).mean(axis=1), axis=1))
        got = y[:B * n * T].reshape(B, n, T)
        assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1)


@pytest.mark.parametrize('T,c', [(64, 2), (128, 1), (40, 2)])
def test_freq_scatter_compiled(emu, T, c):
    """The compile-time freq_scatter (FP32, N = 88, outputs of 88, 44 and 22,
    32 t-values per workgroup with a partial last one at T = 40) against
    numpy FP64 to FP32 rounding."""
    f = emu.lib.kym_emu_freq_scatter
    f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_longlong] + [ctypes.c_int] * 5
    f.restype = ctypes.c_int
    B, A, N, ks = 2, 10, 88, [1, 2, 4, 4, 1]
    rng = np.random.default_rng(T + c)
    x = rng.standard_normal((B, A, T)) + (1j * rng.standard_normal((B, A, T)) if c == 2 else 0)
    xin = (np.stack([x.real, x.imag], -1) if c == 2 else x.real).reshape(-1).astype(np.float32)
    filts = [rng.standard_normal(N).astype(np.float32) for _ in ks]
    ys = [_guarded(B * (N // k) * T, np.float32) for k in ks]
    yp = (ctypes.c_void_p * len(ks))(*[y.ctypes.data for y in ys])
    fp = (ctypes.c_void_p * len(ks))(*[g.ctypes.data for g in filts])
    kp = (ctypes.c_int * len(ks))(*ks)
    assert f(xin.ctypes.data, yp, fp, kp, len(ks), B, A, T, c, N, 0) == 0
    xp = np.zeros((B, N, T), complex)
    xp[:, :A] = x.astype(np.complex64)
    X = np.fft.fft(xp, axis=1)
    for y, g, k in zip(ys, filts, ks):
        n = N // k
        assert np.all(y[B * n * T:] == 12345.0), 'write past the output'
        ref = np.abs(np.fft.ifft((X * g.astype(np.float64)[None, :, None]).reshape(B, k, n, T).mean(axis=1), axis=1))
        got = y[:B * n * T].reshape(B, n, T)
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
