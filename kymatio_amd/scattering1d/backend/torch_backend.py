"""1D HIP backend namespace (plugin name 'torch', the default): op names and
checks of kymatio/scattering1d/backend/torch_backend.py:18-148, each op one
HIP kernel on GPU tensors.  The frontend runs the fused pipeline."""
import ctypes

import torch

from ... import _lib
from ...backend.torch_backend import TorchBackend, _dt, _gpu, _ptr


def _fft1(x, kind, out_last):
    _gpu(x)
    n = x.shape[-2]
    y = torch.empty(x.shape[:-1] + (out_last,), dtype=x.dtype, device=x.device)
    with torch.cuda.device(x.device):
        # 1D transforms run in one pass (W = 1): no scratch needed
        _lib.check(_lib.lib().kym_op_fft(_ptr(x), _ptr(y), None, x.numel() // (n * x.shape[-1]), n, 1, kind, _dt(x),
                                          0, _lib.stream_of(x.device)))
    return y


class TorchBackend1D(TorchBackend):
    @classmethod
    def subsample_fourier(cls, x, k):
        cls.complex_check(x)
        _gpu(x)
        n = x.shape[-2]
        xx = x.contiguous()
        y = torch.empty(x.shape[:-2] + (n // k, 2), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_subsample(_ptr(xx), _ptr(y), x.numel() // (2 * n), n, 1, k, 1, _dt(x),
                                                    _lib.stream_of(x.device)))
        return y

    @staticmethod
    def pad(x, pad_left, pad_right, mode='reflect'):
        if mode != 'constant':
            if pad_left >= x.shape[-1] or pad_right >= x.shape[-1]:
                raise ValueError('Indefinite padding size (larger than tensor).')
        if mode != 'reflect':
            raise NotImplementedError('only reflect padding runs on the HIP path')
        _gpu(x)
        N = x.shape[-1]
        xx = x.contiguous()
        y = torch.empty(x.shape[:-1] + (N + pad_left + pad_right,), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_pad1d(_ptr(xx), _ptr(y), x.numel() // N, N, pad_left, pad_right, _dt(x),
                                                _lib.stream_of(x.device)))
        return y[..., None]

    @staticmethod
    def unpad(x, i0, i1):
        x = x.reshape(x.shape[:-1])
        return x[..., i0:i1]

    @classmethod
    def cfft(cls, x):
        cls.contiguous_check(x)
        cls.complex_check(x)
        return _fft1(x, 0, 2)

    @classmethod
    def rfft(cls, x):
        cls.contiguous_check(x)
        cls.real_check(x)
        return _fft1(x, 2, 2)

    @classmethod
    def irfft(cls, x):
        cls.contiguous_check(x)
        cls.complex_check(x)
        return _fft1(x, 3, 1)

    @classmethod
    def ifft(cls, x):
        cls.contiguous_check(x)
        cls.complex_check(x)
        return _fft1(x, 1, 2)

    @classmethod
    def average_global(cls, x):
        """Sum over time, keeping the axis ((..., T, 1) -> (..., 1, 1)), as
        torch.sum(x, -2, keepdims=True) of the reference; one HIP kernel
        (kym_op_integrals with the power 1, FP64 accumulation)."""
        cls.contiguous_check(x)
        cls.real_check(x)
        _gpu(x)
        T = x.shape[-2]
        rows = x.numel() // max(T, 1)
        out = torch.empty(x.shape[:-2] + (1, 1), dtype=x.dtype, device=x.device)
        if rows and T:
            one = (ctypes.c_double * 1)(1.0)
            with torch.cuda.device(x.device):
                _lib.check(_lib.lib().kym_op_integrals(_ptr(x), _ptr(out), rows, T, one, 1, _dt(x),
                                                        _lib.stream_of(x.device)))
        elif rows:
            out.zero_()
        return out


    @classmethod
    def filter_ifft(cls, x, filt, k, mode='complex'):
        """ifft(subsample_fourier(cdgmm(x, filt), k)) in one HIP kernel
        (kym_op_filter_ifft): x (..., N, 2) complex, filt (N, 1) real ->
        (..., N / k, 2) complex, or its real part / modulus as (..., N / k, 1)
        with mode 'real' / 'modulus'."""
        cls.complex_check(x)
        cls.contiguous_check(x)
        _gpu(x)
        N = x.shape[-2]
        if N % k:
            raise RuntimeError('subsampling factor must divide the size')
        code = {'complex': 0, 'real': 1, 'modulus': 2}[mode]
        f = filt.contiguous()
        if f.dtype != x.dtype or f.numel() != N:
            raise TypeError('Input and filter must be of the same dtype and length.')
        y = torch.empty(x.shape[:-2] + (N // k, 2 if code == 0 else 1), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_filter_ifft(_ptr(x), _ptr(f), _ptr(y), x.numel() // (2 * N), N, k, code,
                                                      _dt(x), _lib.stream_of(x.device)))
        return y

    @classmethod
    def filter_ifft_bank(cls, x, bank, k, mode='complex'):
        """filter_ifft of x (..., N, 2) with every filter of bank (G, N):
        (..., G, N / k, 2 | 1), one HIP kernel (kym_op_filter_ifft_g) for
        what the first order does filter by filter."""
        cls.complex_check(x)
        cls.contiguous_check(x)
        _gpu(x)
        N, G = x.shape[-2], bank.shape[0]
        if N % k:
            raise RuntimeError('subsampling factor must divide the size')
        code = {'complex': 0, 'real': 1, 'modulus': 2}[mode]
        f = bank.contiguous()
        if f.dtype != x.dtype or f.numel() != G * N:
            raise TypeError('Input and filters must be of the same dtype and length.')
        y = torch.empty(x.shape[:-2] + (G, N // k, 2 if code == 0 else 1), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_filter_ifft_g(_ptr(x), _ptr(f), _ptr(y), (x.numel() // (2 * N)) * G, N, k,
                                                        code, G, 0, 0, 0, _dt(x), _lib.stream_of(x.device)))
        return y

    @classmethod
    def filter_ifft_into(cls, x, filt, k, mode, out, goff, gtot):
        """filter_ifft of x (..., G, N, 2) written into out (..., gtot, N / k,
        2 | 1) at [..., goff:goff + G, :, :] (kym_op_filter_ifft_g's row
        map): groups of n1 stacked as backend.stack would, without the copy.
        out None: a new output."""
        cls.complex_check(x)
        cls.contiguous_check(x)
        _gpu(x)
        G, N = x.shape[-3], x.shape[-2]
        if N % k:
            raise RuntimeError('subsampling factor must divide the size')
        code = {'complex': 0, 'real': 1, 'modulus': 2}[mode]
        shape = x.shape[:-3] + (gtot, N // k, 2 if code == 0 else 1)
        if out is None:
            out = torch.empty(shape, dtype=x.dtype, device=x.device)
        if tuple(out.shape) != tuple(shape) or not out.is_contiguous() or goff < 0 or goff + G > gtot:
            raise RuntimeError('output does not fit the rows')
        f = filt.contiguous()
        if f.dtype != x.dtype or f.numel() != N:
            raise TypeError('Input and filter must be of the same dtype and length.')
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_filter_ifft_g(_ptr(x), _ptr(f), _ptr(out), x.numel() // (2 * N), N, k,
                                                        code, 1, G, gtot, goff, _dt(x), _lib.stream_of(x.device)))
        return out

    @classmethod
    def filter_ifft_swap(cls, x, filt, k, mode='modulus'):
        """filter_ifft along the frequency axis of x (batch, T, N, 2), then
        the time and frequency axes swapped: (batch, N / k, T, 2 | 1) -- the
        ifft, swap_time_frequency and (mode 'modulus') modulus of
        frequency scattering in one HIP kernel (kym_op_filter_ifft_t)."""
        cls.complex_check(x)
        cls.contiguous_check(x)
        _gpu(x)
        T, N = x.shape[-3], x.shape[-2]
        if N % k:
            raise RuntimeError('subsampling factor must divide the size')
        code = {'complex': 0, 'real': 1, 'modulus': 2}[mode]
        f = filt.contiguous()
        if f.dtype != x.dtype or f.numel() != N:
            raise TypeError('Input and filter must be of the same dtype and length.')
        y = torch.empty(x.shape[:-3] + (N // k, T, 2 if code == 0 else 1), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_filter_ifft_t(_ptr(x), _ptr(f), _ptr(y), x.numel() // (2 * N), N, k, code,
                                                        T, _dt(x), _lib.stream_of(x.device)))
        return y

    @classmethod
    def freq_scatter(cls, x, filts, ks, N):
        """Frequency scattering of x (..., A, T, 1 | 2) along A (zero-padded to
        N) by every wavelet in filts (each (N, 1)), subsampled by ks, with
        the modulus: [(..., N / k, T, 1)], one HIP kernel per 16 wavelets
        (kym_op_freq_scatter) reading x once."""
        cls.contiguous_check(x)
        _gpu(x)
        c = x.shape[-1]
        if c not in (1, 2):
            raise TypeError('The input should be real or complex.')
        A, T = x.shape[-3], x.shape[-2]
        if A > N:
            raise RuntimeError('the frequency axis is longer than its padded length')
        B = x.numel() // (A * T * c)
        fs = []
        for f in filts:
            f = f.contiguous()
            if f.dtype != x.dtype or f.numel() != N:
                raise TypeError('Input and filter must be of the same dtype and length.')
            fs.append(f)
        for k in ks:
            if N % k:
                raise RuntimeError('subsampling factor must divide the size')
        ys = [torch.empty(x.shape[:-3] + (N // k, T, 1), dtype=x.dtype, device=x.device) for k in ks]
        with torch.cuda.device(x.device):
            for i0 in range(0, len(fs), 16):
                n = min(16, len(fs) - i0)
                yp = (ctypes.c_void_p * n)(*[y.data_ptr() for y in ys[i0:i0 + n]])
                fp = (ctypes.c_void_p * n)(*[f.data_ptr() for f in fs[i0:i0 + n]])
                kp = (ctypes.c_int * n)(*ks[i0:i0 + n])
                _lib.check(_lib.lib().kym_op_freq_scatter(_ptr(x), yp, fp, kp, n, B, A, T, c, N, _dt(x),
                                                           _lib.stream_of(x.device)))
        return ys

    @classmethod
    def lowpass(cls, x, filt, k, axis=-2):
        """irfft(subsample_fourier(cdgmm(rfft(.), filt), k)) along the time
        axis (-2) or the frequency axis (-3) of a real x (..., A, T, 1), in
        one HIP kernel (kym_op_lowpass); that axis shrinks by k.  Along the
        frequency axis it is what swap_time_frequency, rfft, cdgmm,
        subsample_fourier, irfft and swap_time_frequency compute."""
        cls.real_check(x)
        cls.contiguous_check(x)
        _gpu(x)
        if axis == -2:
            N, S = x.shape[-2], 1
            shape = x.shape[:-2] + (N // k, 1)
        elif axis == -3:
            N, S = x.shape[-3], x.shape[-2]
            shape = x.shape[:-3] + (N // k, S, 1)
        else:
            raise ValueError('axis must be -2 (time) or -3 (frequency)')
        if N % k:
            raise RuntimeError('subsampling factor must divide the size')
        f = filt.contiguous()
        if f.dtype != x.dtype or f.numel() != N:
            raise TypeError('Input and filter must be of the same dtype and length.')
        y = torch.empty(shape, dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_lowpass(_ptr(x), _ptr(f), _ptr(y), x.numel() // (N * S), N, S, k, _dt(x),
                                                  _lib.stream_of(x.device)))
        return y

    # ---- frequency-axis layout ops of time-frequency scattering
    # (kymatio/scattering1d/backend/torch_backend.py:150-223); the axis n1
    # sits before time, tensors end in a real/imag (or singleton) axis.
    @classmethod
    def pad_frequency(cls, x, padding):
        """Zeros appended to the n1 axis (one-sided: towards lower frequencies)."""
        return torch.nn.functional.pad(x, (0, 0, 0, padding), mode='constant', value=0)

    @classmethod
    def swap_time_frequency(cls, x):
        """(batch, a, b, 1|2) -> (batch, b, a, 1|2), contiguous."""
        return torch.transpose(x, dim0=-2, dim1=-3).contiguous()

    @staticmethod
    def unpad_frequency(x, n1_max, n1_stride):
        return x[:, :, :1 + (n1_max // n1_stride), :]

    @staticmethod
    def split_frequency_axis(x):
        return torch.split(x, 1, dim=-3)


backend = TorchBackend1D
