"""Scattering1D for NumPy arrays on the HIP path
(kymatio/scattering1d/frontend/numpy_frontend.py:7-15 with the shared
checks of kymatio/scattering1d/frontend/base_frontend.py:106-155): same
constructor, meta(), output_size(), out_type 'array' | 'list' | 'dict';
the arrays are staged through the device
(kymatio_amd/frontend/numpy_frontend.py)."""
import torch

from ...frontend.numpy_frontend import ScatteringNumPy
from .jtfs_frontend import TimeFrequencyScatteringTorch1D
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


class TimeFrequencyScatteringNumPy(ScatteringNumPy1D):
    """TimeFrequencyScattering for NumPy arrays
    (kymatio/scattering1d/frontend/numpy_frontend.py:18-34); meta() from the
    CPU-side module, the transform staged through the device."""

    def __init__(self, J, J_fr, Q, shape, T=None, stride=None, Q_fr=1, F=None, stride_fr=None, out_type='array',
                 format='time', backend='numpy', device=None):
        ScatteringNumPy.__init__(self)
        if isinstance(backend, str):
            if not backend.startswith('numpy'):
                raise ImportError('The backend ' + backend + ' is not supported by the numpy frontend.')
        elif not str(getattr(backend, 'name', '')).startswith('numpy'):
            raise ImportError('The backend ' + str(getattr(backend, 'name', backend)) + ' is not supported by the '
                              'numpy frontend.')
        self.J, self.J_fr, self.Q, self.shape, self.T, self.stride = J, J_fr, Q, shape, T, stride
        self.Q_fr, self.F, self.stride_fr, self.out_type, self.format = Q_fr, F, stride_fr, out_type, format
        self.backend, self.device = backend, device
        self._torch_modules = {}
        self._spec = self._make_torch_module()

    def _make_torch_module(self):
        return TimeFrequencyScatteringTorch1D(self.J, self.J_fr, self.Q, self.shape, T=self.T, stride=self.stride,
                                              Q_fr=self.Q_fr, F=self.F, stride_fr=self.stride_fr,
                                              out_type=self.out_type, format=self.format)

    def output_size(self, detail=False):
        raise NotImplementedError('output_size is not defined for TimeFrequencyScattering; use meta().')


__all__ = ['ScatteringNumPy1D', 'TimeFrequencyScatteringNumPy']
