"""kymatio_amd.numpy: the scattering transforms for NumPy arrays on MI355X
(drop-in names of kymatio/numpy.py:1-27; arrays staged through the device)."""
from .scattering1d.frontend.numpy_frontend import ScatteringNumPy1D as Scattering1D
from .scattering1d.frontend.numpy_frontend import TimeFrequencyScatteringNumPy as TimeFrequencyScattering
from .scattering2d.frontend.numpy_frontend import ScatteringNumPy2D as Scattering2D
from .scattering3d.frontend.numpy_frontend import HarmonicScatteringNumPy3D as HarmonicScattering3D

Scattering1D.__module__ = 'kymatio_amd.numpy'
Scattering1D.__name__ = 'Scattering1D'
Scattering2D.__module__ = 'kymatio_amd.numpy'
Scattering2D.__name__ = 'Scattering2D'

TimeFrequencyScattering.__module__ = 'kymatio_amd.numpy'
TimeFrequencyScattering.__name__ = 'TimeFrequencyScattering'

HarmonicScattering3D.__module__ = 'kymatio_amd.numpy'
HarmonicScattering3D.__name__ = 'HarmonicScattering3D'

__all__ = ['Scattering1D', 'Scattering2D', 'TimeFrequencyScattering', 'HarmonicScattering3D']
