"""HarmonicScattering3D as a torch.nn.Module over batched HIP launches.

Interface of kymatio/scattering3d/frontend/torch_frontend.py:7-80 (with
kymatio/scattering3d/frontend/base_frontend.py:4-36): the constructor
arguments, the buffer names (tensor{l}: the order-l wavelets as
(J + 1, 2l + 1, M, N, O, 2) real/imaginary pairs; tensor_gaussian_filter:
the Gaussians, imaginary part zero), the checks with their messages and
the output layout batch_shape + (paths, L + 1, P).  The transform itself is
kymatio_amd/scattering3d/core/harmonic.py, three launches per (l, j) over
every m and image.
"""
import numpy as np
import torch

from ...frontend.base_frontend import ScatteringBase
from ...frontend.torch_frontend import ScatteringTorch
from ..core.harmonic import harmonic_scattering
from ..filter_bank import gaussian_filter_bank, solid_harmonic_filter_bank

_METHODS = ('integral',)


def _as_pairs(a, keep_imag=True):
    """complex ndarray -> float32 tensor of (re, im) pairs in a last axis."""
    c = np.asarray(a)
    re = torch.from_numpy(np.ascontiguousarray(c.real, dtype=np.float32))
    im = torch.from_numpy(np.ascontiguousarray(c.imag, dtype=np.float32)) if keep_imag else torch.zeros_like(re)
    return torch.stack((re, im), dim=-1)


class HarmonicScatteringTorch3D(ScatteringTorch, ScatteringBase):
    def __init__(self, J, shape, L=3, sigma_0=1, max_order=2, rotation_covariant=True, method='integral', points=None,
                 integral_powers=(0.5, 1., 2.), backend='torch'):
        ScatteringTorch.__init__(self)
        self.J, self.shape, self.L, self.sigma_0 = J, shape, L, sigma_0
        self.max_order, self.rotation_covariant, self.method = max_order, rotation_covariant, method
        self.points, self.integral_powers, self.backend = points, integral_powers, backend
        self.build()

    def build(self):
        self._instantiate_backend('kymatio_amd.scattering3d.backend.')
        self.M, self.N, self.O = self.shape
        bank = solid_harmonic_filter_bank(self.M, self.N, self.O, self.J, self.L, self.sigma_0)
        for l, psi_l in enumerate(bank):
            self.register_buffer('tensor%d' % l, _as_pairs(psi_l))
        self.register_buffer('tensor_gaussian_filter',
                             _as_pairs(gaussian_filter_bank(self.M, self.N, self.O, self.J + 1, self.sigma_0),
                                       keep_imag=False))

    def register_filters(self):
        """The buffers are registered by build()."""

    @property
    def filters(self):
        """Per l, the (J + 1, 2l + 1, M, N, O, 2) wavelet buffer on the module's device."""
        return [getattr(self, 'tensor%d' % l) for l in range(self.L + 1)]

    @property
    def gaussian_filters(self):
        return self.tensor_gaussian_filter

    def _check(self, x):
        if not torch.is_tensor(x):
            raise TypeError('The input should be a torch.cuda.FloatTensor, a torch.FloatTensor or a '
                            'torch.DoubleTensor.')
        if x.dim() < 3:
            raise RuntimeError('Input tensor must have at least three dimensions.')
        if tuple(x.shape[-3:]) != (self.M, self.N, self.O):
            raise RuntimeError('Tensor must be of spatial size (%i, %i, %i).' % (self.M, self.N, self.O))
        if self.method not in _METHODS:
            raise ValueError('method must be in {}'.format(list(_METHODS)))
        if not x.is_cuda:
            raise TypeError('HarmonicScattering3D runs on the GPU (HIP); got a CPU tensor.')

    def scattering(self, input_array):
        self._check(input_array)
        lead = input_array.shape[:-3]
        volumes = input_array.contiguous().reshape((-1, self.M, self.N, self.O, 1))
        S = harmonic_scattering(volumes, self.filters, self.J, self.L, self.max_order, self.rotation_covariant,
                                self.integral_powers, self.backend)
        return S.reshape(tuple(lead) + tuple(S.shape[1:]))


__all__ = ['HarmonicScatteringTorch3D']
