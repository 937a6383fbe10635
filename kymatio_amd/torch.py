"""kymatio_amd.torch: the torch.nn.Module scattering transforms on MI355X
(drop-in names of kymatio/torch.py:1-27)."""
from .scattering1d.frontend.jtfs_frontend import TimeFrequencyScatteringTorch1D as TimeFrequencyScattering
from .scattering1d.frontend.torch_frontend import ScatteringTorch1D as Scattering1D
from .scattering2d.frontend.torch_frontend import ScatteringTorch2D as Scattering2D
from .scattering3d.frontend.torch_frontend import HarmonicScatteringTorch3D as HarmonicScattering3D

Scattering1D.__module__ = 'kymatio_amd.torch'
Scattering1D.__name__ = 'Scattering1D'
Scattering2D.__module__ = 'kymatio_amd.torch'
Scattering2D.__name__ = 'Scattering2D'

TimeFrequencyScattering.__module__ = 'kymatio_amd.torch'
TimeFrequencyScattering.__name__ = 'TimeFrequencyScattering'

HarmonicScattering3D.__module__ = 'kymatio_amd.torch'
HarmonicScattering3D.__name__ = 'HarmonicScattering3D'

__all__ = ['Scattering1D', 'Scattering2D', 'TimeFrequencyScattering', 'HarmonicScattering3D']
