"""HarmonicScattering3D as a scikit-learn transformer
(kymatio/scattering3d/frontend/sklearn_frontend.py)."""
from ...frontend.sklearn_frontend import ScatteringTransformerMixin
from .numpy_frontend import HarmonicScatteringNumPy3D


class HarmonicScatteringTransformer3D(ScatteringTransformerMixin, HarmonicScatteringNumPy3D):
    pass


__all__ = ['HarmonicScatteringTransformer3D']
