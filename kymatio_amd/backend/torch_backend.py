"""Backend namespace shared by the 1D and 2D HIP backends: the checks,
helpers and element-wise ops of kymatio/backend/torch_backend.py:99-232 with
the same messages; the ops run the per-op HIP kernels (kym_op_* in
include/kymatio_hip.h) on GPU tensors (forward only)."""
import ctypes

import torch

from .. import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _dt(t):
    if t.dtype not in _lib.DTYPE_CODE:
        raise TypeError('Input should be float32 or float64.')
    return _lib.DTYPE_CODE[t.dtype]


def _gpu(t):
    if not t.is_cuda:
        raise TypeError('Use the torch backend (without HIP) for CPU tensors: the HIP ops need GPU tensors.')


class TorchBackend:
    name = 'torch'
    # the frontend runs its fused HIP pipeline for backends flagged `fused`
    fused = True

    @classmethod
    def input_checks(cls, x):
        if x is None:
            raise TypeError('The input should be not empty.')
        cls.contiguous_check(x)

    @classmethod
    def complex_check(cls, x):
        if not cls._is_complex(x):
            raise TypeError('The input should be complex (i.e. last dimension is 2).')

    @classmethod
    def real_check(cls, x):
        if not cls._is_real(x):
            raise TypeError('The input should be real.')

    @classmethod
    def complex_contiguous_check(cls, x):
        cls.complex_check(x)
        cls.contiguous_check(x)

    @staticmethod
    def contiguous_check(x):
        if not x.is_contiguous():
            raise RuntimeError('Tensors must be contiguous.')

    @staticmethod
    def _is_complex(x):
        return x.shape[-1] == 2

    @staticmethod
    def _is_real(x):
        return x.shape[-1] == 1

    @staticmethod
    def stack(arrays, dim=2):
        return torch.stack(arrays, dim=dim)

    @classmethod
    def modulus(cls, x):
        """|x| of a complex (..., 2) tensor as a real (..., 1) tensor
        (kymatio/backend/torch_backend.py:137-141)."""
        cls.complex_contiguous_check(x)
        _gpu(x)
        y = torch.empty(x.shape[:-1] + (1,), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_modulus(_ptr(x), _ptr(y), x.numel() // 2, _dt(x),
                                                  _lib.stream_of(x.device)))
        return y

    @classmethod
    def cdgmm(cls, A, B):
        """Complex point-wise product A * B broadcast over the leading axes of A
        (kymatio/backend/torch_backend.py:147-219), same checks and messages."""
        if not cls._is_real(B):
            cls.complex_contiguous_check(B)
        else:
            cls.contiguous_check(B)
        cls.complex_contiguous_check(A)
        if A.shape[-len(B.shape):-1] != B.shape[:-1]:
            raise RuntimeError('The filters are not compatible for multiplication.')
        if A.dtype is not B.dtype:
            raise TypeError('Input and filter must be of the same dtype.')
        if B.device.type == 'cuda':
            if A.device.type == 'cuda':
                if A.device.index != B.device.index:
                    raise TypeError('Input and filter must be on the same GPU.')
            else:
                raise TypeError('Input must be on GPU.')
        if B.device.type == 'cpu':
            if A.device.type == 'cuda':
                raise TypeError('Input must be on CPU.')
        _gpu(A)
        C = torch.empty_like(A)
        n = B.numel() // B.shape[-1]
        with torch.cuda.device(A.device):
            _lib.check(_lib.lib().kym_op_cdgmm(_ptr(A), _ptr(B), _ptr(C), A.numel() // (2 * n), n,
                                                int(not cls._is_real(B)), _dt(A), _lib.stream_of(A.device)))
        return C

    @staticmethod
    def reshape_input(x, signal_shape):
        return x.reshape((-1, 1) + signal_shape)

    @staticmethod
    def reshape_output(S, batch_shape, n_kept_dims):
        return S.reshape(batch_shape + S.shape[-n_kept_dims:])

    @staticmethod
    def shape(x):
        return x.shape


backend = TorchBackend
