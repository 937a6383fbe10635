"""Scattering1D as a scikit-learn transformer
(kymatio/scattering1d/frontend/sklearn_frontend.py:1-22)."""
from ...frontend.sklearn_frontend import ScatteringTransformerMixin
from .numpy_frontend import ScatteringNumPy1D, TimeFrequencyScatteringNumPy


class ScatteringTransformer1D(ScatteringTransformerMixin, ScatteringNumPy1D):
    pass


class TimeFrequencyScatteringTransformer(ScatteringTransformerMixin, TimeFrequencyScatteringNumPy):
    pass


__all__ = ['ScatteringTransformer1D', 'TimeFrequencyScatteringTransformer']
