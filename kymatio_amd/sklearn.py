"""kymatio_amd.sklearn: the scattering transforms as scikit-learn
transformers on MI355X (drop-in names of kymatio/sklearn.py:1-12)."""
from .scattering1d.frontend.sklearn_frontend import ScatteringTransformer1D as Scattering1D
from .scattering1d.frontend.sklearn_frontend import TimeFrequencyScatteringTransformer as TimeFrequencyScattering
from .scattering2d.frontend.sklearn_frontend import ScatteringTransformer2D as Scattering2D
from .scattering3d.frontend.sklearn_frontend import HarmonicScatteringTransformer3D as HarmonicScattering3D

__all__ = ['Scattering1D', 'Scattering2D', 'TimeFrequencyScattering', 'HarmonicScattering3D']
