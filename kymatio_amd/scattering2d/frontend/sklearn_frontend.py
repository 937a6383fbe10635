"""Scattering2D as a scikit-learn transformer
(kymatio/scattering2d/frontend/sklearn_frontend.py:1-14)."""
from ...frontend.sklearn_frontend import ScatteringTransformerMixin
from .numpy_frontend import ScatteringNumPy2D


class ScatteringTransformer2D(ScatteringTransformerMixin, ScatteringNumPy2D):
    pass


__all__ = ['ScatteringTransformer2D']
