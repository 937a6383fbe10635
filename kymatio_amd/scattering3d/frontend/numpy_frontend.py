"""HarmonicScattering3D for NumPy arrays on the HIP path
(kymatio/scattering3d/frontend/numpy_frontend.py): the reference's checks,
the array staged through the device (kymatio_amd/frontend/numpy_frontend.py)."""
import numpy as np
import torch

from ...frontend.numpy_frontend import ScatteringNumPy
from .torch_frontend import HarmonicScatteringTorch3D


class HarmonicScatteringNumPy3D(ScatteringNumPy):
    def __init__(self, J, shape, L=3, sigma_0=1, max_order=2, rotation_covariant=True, method='integral', points=None,
                 integral_powers=(0.5, 1., 2.), backend='numpy', device=None):
        ScatteringNumPy.__init__(self)
        if isinstance(backend, str):
            if not backend.startswith('numpy'):
                raise ImportError('The backend ' + backend + ' is not supported by the numpy frontend.')
        elif not str(getattr(backend, 'name', '')).startswith('numpy'):
            raise ImportError('The backend ' + str(getattr(backend, 'name', backend)) + ' is not supported by the '
                              'numpy frontend.')
        self.J, self.shape, self.L, self.sigma_0, self.max_order = J, shape, L, sigma_0, max_order
        self.rotation_covariant, self.method, self.points = rotation_covariant, method, points
        self.integral_powers, self.backend, self.device = integral_powers, backend, device
        self._torch_modules = {}

    def _make_torch_module(self):
        return HarmonicScatteringTorch3D(self.J, self.shape, L=self.L, sigma_0=self.sigma_0,
                                         max_order=self.max_order, rotation_covariant=self.rotation_covariant,
                                         method=self.method, points=self.points,
                                         integral_powers=self.integral_powers)

    def scattering(self, input_array):
        self.input_checks(input_array)
        if input_array.ndim < 3:
            raise RuntimeError('Input array must have at least three dimensions.')
        if tuple(input_array.shape[-3:]) != tuple(self.shape):
            raise RuntimeError('Array must be of spatial size (%i, %i, %i).' % tuple(self.shape))
        x, dtype = self._stage_in(input_array)
        m = self._module(dtype)
        m.max_order, m.method, m.integral_powers = self.max_order, self.method, self.integral_powers
        with torch.no_grad():
            S = m.scattering(x)
        return self._stage_out(S)


__all__ = ['HarmonicScatteringNumPy3D']
