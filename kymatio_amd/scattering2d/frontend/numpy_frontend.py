"""Scattering2D for NumPy arrays on the HIP path
(kymatio/scattering2d/frontend/numpy_frontend.py:1-53): same constructor,
checks, messages and output layout; the arrays are staged through the
device (kymatio_amd/frontend/numpy_frontend.py)."""
import torch

from ...frontend.numpy_frontend import ScatteringNumPy
from ..utils import compute_padding
from .torch_frontend import ScatteringTorch2D


class ScatteringNumPy2D(ScatteringNumPy):
    def __init__(self, J, shape, L=8, max_order=2, pre_pad=False, backend='numpy', out_type='array', device=None):
        ScatteringNumPy.__init__(self)
        self.J, self.shape, self.L, self.max_order = J, tuple(shape), L, max_order
        self.pre_pad, self.backend, self.out_type, self.device = pre_pad, backend, out_type, device
        if isinstance(backend, str):
            if not backend.startswith('numpy'):
                raise ImportError('The backend ' + backend + ' is not supported by the numpy frontend.')
        elif not str(getattr(backend, 'name', '')).startswith('numpy'):
            raise ImportError('The backend ' + str(getattr(backend, 'name', backend)) + ' is not supported by the '
                              'numpy frontend.')
        M, N = self.shape
        if 2 ** J > M or 2 ** J > N:
            raise RuntimeError('The smallest dimension should be larger than 2^J.')
        self._M_padded, self._N_padded = compute_padding(M, N, J)
        self._torch_modules = {}

    def _make_torch_module(self):
        return ScatteringTorch2D(self.J, self.shape, L=self.L, max_order=self.max_order, pre_pad=self.pre_pad,
                                 out_type=self.out_type)

    def scattering(self, input):
        self.input_checks(input)
        if len(input.shape) < 2:
            raise RuntimeError('Input array must have at least two dimensions.')
        if (input.shape[-1] != self.shape[-1] or input.shape[-2] != self.shape[-2]) and not self.pre_pad:
            raise RuntimeError('NumPy array must be of spatial size (%i,%i).' % (self.shape[0], self.shape[1]))
        if (input.shape[-1] != self._N_padded or input.shape[-2] != self._M_padded) and self.pre_pad:
            raise RuntimeError('Padded array must be of spatial size (%i,%i).' % (self._M_padded, self._N_padded))
        if self.out_type not in ('array', 'list'):
            raise RuntimeError("The out_type must be one of 'array' or 'list'.")
        x, dtype = self._stage_in(input)
        m = self._module(dtype)
        m.out_type = self.out_type
        with torch.no_grad():
            S = m.scattering(x)
        return self._stage_out(S)


__all__ = ['ScatteringNumPy2D']
