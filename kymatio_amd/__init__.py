"""kymatio_amd -- MI355X-native scattering transforms (Scattering2D, Scattering1D).

A drop-in for the torch frontend of kymatio: ``kymatio_amd.torch.Scattering2D``
and ``kymatio_amd.torch.Scattering1D`` keep the reference's constructor
arguments, buffers, output layout and exceptions, and run hand-written HIP
kernels for gfx950 through the C ABI in include/kymatio_hip.h.
"""
from .scattering1d import ScatteringEntry1D as Scattering1D
from .scattering2d import ScatteringEntry2D as Scattering2D
from .scattering3d import HarmonicScatteringEntry3D as HarmonicScattering3D

__version__ = '0.1.0'
__all__ = ['Scattering1D', 'Scattering2D', 'HarmonicScattering3D']
