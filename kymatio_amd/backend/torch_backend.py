"""Backend namespace shared by the 1D and 2D HIP backends: the checks and
helpers of kymatio/backend/torch_backend.py:99-232 with the same messages.
Compute ops live in the dimension-specific modules."""
import torch


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
