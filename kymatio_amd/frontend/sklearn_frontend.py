"""scikit-learn transformers over the NumPy frontends.

Interface of kymatio/frontend/sklearn_frontend.py:1-45: estimators whose
parameters are the constructor arguments, `fit` learning nothing, and
`transform` / `predict` mapping a 2D array of flattened signals
(n_samples, prod(shape)) to a 2D array of flattened coefficients
(n_samples, n_coefficients), so a scattering transform can sit in a
sklearn Pipeline.  A row may hold several signals (a multiple of
prod(shape) values); its coefficients are then concatenated, and leading
axes are kept.
"""
import numpy as np
from sklearn.base import BaseEstimator, TransformerMixin


class ScatteringTransformerMixin(BaseEstimator, TransformerMixin):
    def fit(self, x=None, y=None):
        """The transform has no state to learn; returns the estimator."""
        return self

    def _signal_shape(self):
        s = self.shape
        return tuple(int(v) for v in s) if isinstance(s, (tuple, list, np.ndarray)) else (int(s),)

    def transform(self, x):
        """Rows of flattened signals -> rows of flattened coefficients."""
        rows = tuple(x.shape[:-1])
        coeffs = self.scattering(x.reshape((-1,) + self._signal_shape()))
        return coeffs.reshape(rows + (-1,))

    def predict(self, x):
        """Same as transform (the reference exposes both names)."""
        return self.transform(x)
