"""2D HIP backend namespace (plugin name 'torch', the default): the op
names and checks of kymatio/scattering2d/backend/torch_backend.py:19-183,
each op one HIP kernel on GPU tensors (kym_op_*).  The frontend runs the
fused pipeline for this backend; ``torch_ops`` runs the reference core loop
over these ops instead."""
import torch

from ... import _lib
from ...backend.torch_backend import TorchBackend, _dt, _gpu, _ptr


class Pad:
    def __init__(self, pad_size, input_size):
        """Reflect padding [top, bottom, left, right] of an input of size
        input_size, output with a trailing real axis."""
        self.pad_size = list(pad_size)
        self.input_size = list(input_size)

    def __call__(self, x):
        _gpu(x)
        batch_shape, (M, N) = x.shape[:-2], x.shape[-2:]
        t, b, l, r = self.pad_size
        xx = x.reshape((-1, M, N)).contiguous()
        y = torch.empty((xx.shape[0], M + t + b, N + l + r), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_pad2d(_ptr(xx), _ptr(y), xx.shape[0], M, N, t, b, l, r, _dt(x),
                                                _lib.stream_of(x.device)))
        return y.reshape(batch_shape + y.shape[-2:] + (1,))


def _fft(x, kind, out_last):
    _gpu(x)
    H, W = x.shape[-3], x.shape[-2]
    batch = x.numel() // (H * W * x.shape[-1])
    y = torch.empty(x.shape[:-1] + (out_last,), dtype=x.dtype, device=x.device)
    # complex scratch for a real output computed in two passes
    work = torch.empty(x.shape[:-1] + (2,), dtype=x.dtype, device=x.device) if kind == 3 else None
    with torch.cuda.device(x.device):
        _lib.check(_lib.lib().kym_op_fft(_ptr(x), _ptr(y), _ptr(work) if work is not None else None, batch, H, W,
                                          kind, _dt(x), 0, _lib.stream_of(x.device)))
    return y


class TorchBackend2D(TorchBackend):
    Pad = Pad

    @classmethod
    def subsample_fourier(cls, x, k):
        cls.contiguous_check(x)
        cls.complex_check(x)
        _gpu(x)
        H, W = x.shape[-3], x.shape[-2]
        y = torch.empty(x.shape[:-3] + (H // k, W // k, 2), dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            _lib.check(_lib.lib().kym_op_subsample(_ptr(x), _ptr(y), x.numel() // (2 * H * W), H, W, k, k, _dt(x),
                                                    _lib.stream_of(x.device)))
        return y

    @classmethod
    def rfft(cls, x):
        cls.contiguous_check(x)
        cls.real_check(x)
        return _fft(x, 2, 2)

    @classmethod
    def irfft(cls, x):
        cls.contiguous_check(x)
        cls.complex_check(x)
        return _fft(x, 3, 1)

    @classmethod
    def ifft(cls, x):
        cls.contiguous_check(x)
        cls.complex_check(x)
        return _fft(x, 1, 2)

    @staticmethod
    def unpad(in_):
        _gpu(in_)
        H, W = in_.shape[-3], in_.shape[-2]
        x = in_.contiguous()
        y = torch.empty(in_.shape[:-3] + (H - 2, W - 2), dtype=in_.dtype, device=in_.device)
        with torch.cuda.device(in_.device):
            _lib.check(_lib.lib().kym_op_unpad2d(_ptr(x), _ptr(y), x.numel() // (H * W * in_.shape[-1]), H, W,
                                                  _dt(in_), _lib.stream_of(in_.device)))
        return y

    @staticmethod
    def stack(arrays):
        return TorchBackend.stack(arrays, -3)


backend = TorchBackend2D
