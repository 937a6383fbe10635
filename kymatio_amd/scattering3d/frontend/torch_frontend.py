"""HarmonicScattering3D as a torch.nn.Module on the HIP ops
(kymatio/scattering3d/frontend/torch_frontend.py:7-80 over
kymatio/scattering3d/frontend/base_frontend.py:4-36): same constructor,
buffers (tensor0.. the solid harmonic filters per l, then
tensor_gaussian_filter), checks, messages and output layout
(batch, paths, L + 1, P)."""
import torch

from ...frontend.base_frontend import ScatteringBase
from ...frontend.torch_frontend import ScatteringTorch
from ..core.scattering3d import scattering3d
from ..filter_bank import gaussian_filter_bank, solid_harmonic_filter_bank


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
        self.filters = solid_harmonic_filter_bank(self.M, self.N, self.O, self.J, self.L, self.sigma_0)
        self.gaussian_filters = gaussian_filter_bank(self.M, self.N, self.O, self.J + 1, self.sigma_0)
        self.register_filters()

    def register_filters(self):
        for k in range(len(self.filters)):
            f = torch.zeros(self.filters[k].shape + (2,))
            f[..., 0] = torch.from_numpy(self.filters[k].real).reshape(self.filters[k].shape)
            f[..., 1] = torch.from_numpy(self.filters[k].imag).reshape(self.filters[k].shape)
            self.filters[k] = f
            self.register_buffer('tensor' + str(k), f)
        g = torch.zeros(self.gaussian_filters.shape + (2,))
        g[..., 0] = torch.from_numpy(self.gaussian_filters.real)
        self.gaussian_filters = g
        self.register_buffer('tensor_gaussian_filter', g)

    def scattering(self, input_array):
        if not torch.is_tensor(input_array):
            raise TypeError('The input should be a torch.cuda.FloatTensor, a torch.FloatTensor or a '
                            'torch.DoubleTensor.')
        if input_array.dim() < 3:
            raise RuntimeError('Input tensor must have at least three dimensions.')
        if input_array.shape[-1] != self.O or input_array.shape[-2] != self.N or input_array.shape[-3] != self.M:
            raise RuntimeError('Tensor must be of spatial size (%i, %i, %i).' % (self.M, self.N, self.O))
        if not input_array.is_cuda:
            raise TypeError('HarmonicScattering3D runs on the GPU (HIP); got a CPU tensor.')
        x = input_array.contiguous()
        batch_shape, signal_shape = x.shape[:-3], x.shape[-3:]
        x = x.reshape((-1,) + signal_shape + (1,))
        bufs = dict(self.named_buffers())
        for k in range(len(self.filters)):
            self.filters[k] = bufs['tensor' + str(k)]
        if self.method not in ['integral']:
            raise ValueError('method must be in {}'.format(['integral']))
        averaging = lambda U: self.backend.compute_integrals(U, self.integral_powers)  # noqa: E731
        S = scattering3d(x, filters=self.filters, rotation_covariant=self.rotation_covariant, L=self.L, J=self.J,
                         max_order=self.max_order, backend=self.backend, averaging=averaging)
        return S.reshape(batch_shape + S.shape[1:])


__all__ = ['HarmonicScatteringTorch3D']
