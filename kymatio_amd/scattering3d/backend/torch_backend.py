"""3D HIP backend namespace (op names and checks of
kymatio/scattering3d/backend/torch_backend.py:59-151): the 3D FFTs are the 2D
HIP transform over the last two axes followed by the 1D HIP transform over
the first (after a transpose); products, moduli and integrals are HIP
kernels."""
import ctypes

import torch

from ... import _lib
from ...backend.torch_backend import TorchBackend, _dt, _gpu, _ptr


def _fft_axis0(x, kind):
    """1D transform over M of (B, M, N, O, 2)."""
    B, M, N, O, _ = x.shape
    xt = x.reshape(B, M, N * O, 2).transpose(1, 2).contiguous()  # (B, N O, M, 2)
    yt = torch.empty_like(xt)
    with torch.cuda.device(x.device):
        _lib.check(_lib.lib().kym_op_fft(_ptr(xt), _ptr(yt), None, B * N * O, M, 1, kind, _dt(x), 0,
                                          _lib.stream_of(x.device)))
    return yt.transpose(1, 2).reshape(B, M, N, O, 2).contiguous()


def _fft3(x, kind):
    """kind 0 forward c2c, 1 inverse c2c (1/(M N O)), 2 forward from real."""
    _gpu(x)
    B, M, N, O = x.shape[:4]
    y = torch.empty((B, M, N, O, 2), dtype=x.dtype, device=x.device)
    with torch.cuda.device(x.device):
        _lib.check(_lib.lib().kym_op_fft(_ptr(x), _ptr(y), None, B * M, N, O, kind, _dt(x), 0,
                                          _lib.stream_of(x.device)))
    return _fft_axis0(y, 1 if kind == 1 else 0)


class TorchBackend3D(TorchBackend):
    @staticmethod
    def stack(arrays, L):
        S = torch.stack(arrays, dim=1)
        return S.reshape((S.shape[0], S.shape[1] // (L + 1), L + 1) + S.shape[2:])

    @classmethod
    def rfft(cls, x):
        cls.contiguous_check(x)
        cls.real_check(x)
        return _fft3(x, 2)

    @classmethod
    def ifft(cls, x):
        cls.contiguous_check(x)
        cls.complex_check(x)
        return _fft3(x, 1)

    @classmethod
    def cdgmm3d(cls, A, B):
        return cls.cdgmm(A, B)

    @staticmethod
    def modulus_rotation(x, module=None):
        """sqrt(module^2 + |x|^2) (the first m: |x|), (..., 2) -> (..., 1)."""
        _gpu(x)
        xx = x.contiguous()
        y = torch.empty(x.shape[:-1] + (1,), dtype=x.dtype, device=x.device)
        prev = None if module is None else module.contiguous()
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_modulus_rotation(_ptr(xx), None if prev is None else _ptr(prev), _ptr(y),
                                                           xx.numel() // 2, _dt(x), _lib.stream_of(x.device)))
        return y

    @staticmethod
    def compute_integrals(input_array, integral_powers):
        """(B, M, N, O[, 1]) -> (B, P): sum of x^p over space for each p."""
        _gpu(input_array)
        x = input_array.contiguous()
        B = x.shape[0]
        P = len(integral_powers)
        pw = (ctypes.c_double * P)(*[float(p) for p in integral_powers])
        out = torch.empty((B, P), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_integrals(_ptr(x), _ptr(out), B, x.numel() // max(B, 1), pw, P, _dt(x),
                                                    _lib.stream_of(x.device)))
        # the reference returns its integrals in torch's default float32
        return out.to(torch.float32)


backend = TorchBackend3D
