"""Scattering1D for NumPy arrays on the HIP path
(kymatio/scattering1d/frontend/numpy_frontend.py:7-15 with the shared
checks of kymatio/scattering1d/frontend/base_frontend.py:106-155): same
constructor, meta(), output_size(), out_type 'array' | 'list' | 'dict';
the arrays are staged through the device
(kymatio_amd/frontend/numpy_frontend.py)."""
import torch

from ...frontend.numpy_frontend import ScatteringNumPy
from .torch_frontend import ScatteringTorch1D


class ScatteringNumPy1D(ScatteringNumPy):
    def __init__(self, J, shape, Q=1, T=None, stride=None, max_order=2, oversampling=0, out_type='array',
                 backend='numpy', device=None):
        ScatteringNumPy.__init__(self)
        if isinstance(backend, str):
            if not backend.startswith('numpy'):
                raise ImportError('The backend ' + backend + ' is not supported by the numpy frontend.')
        elif not str(getattr(backend, 'name', '')).startswith('numpy'):
            raise ImportError('The backend ' + str(getattr(backend, 'name', backend)) + ' is not supported by the '
                              'numpy frontend.')
        self.J, self.shape, self.Q, self.T, self.stride = J, shape, Q, T, stride
        self.max_order, self.oversampling, self.out_type, self.backend, self.device = (
            max_order, oversampling, out_type, backend, device)
        self._torch_modules = {}
        # the CPU-side module: constructor checks, meta() and output_size()
        self._spec = self._make_torch_module()

    def _make_torch_module(self):
        return ScatteringTorch1D(self.J, self.shape, Q=self.Q, T=self.T, stride=self.stride,
                                 max_order=self.max_order, oversampling=self.oversampling, out_type=self.out_type)

    def meta(self):
        return self._spec.meta()

    def output_size(self, detail=False):
        return self._spec.output_size(detail)

    def scattering(self, x):
        self.input_checks(x)
        if len(x.shape) < 1:
            raise ValueError('Input tensor x should have at least one axis, got {}'.format(len(x.shape)))
        xt, dtype = self._stage_in(x)
        m = self._module(dtype)
        m.out_type = self.out_type
        with torch.no_grad():
            S = m.scattering(xt)
        return self._stage_out(S)


__all__ = ['ScatteringNumPy1D']
