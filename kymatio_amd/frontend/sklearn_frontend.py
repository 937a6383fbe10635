"""scikit-learn transformers over the NumPy frontends
(kymatio/frontend/sklearn_frontend.py:1-45): fit is a no-op, predict /
transform take rows of flattened signals and return rows of flattened
scattering coefficients, so the transforms fit in a sklearn Pipeline."""
from sklearn.base import BaseEstimator, TransformerMixin


class ScatteringTransformerMixin(BaseEstimator, TransformerMixin):
    def fit(self, x=None, y=None):
        # No fitting necessary.
        return self

    def predict(self, x):
        batch_shape = x.shape[:-1]
        shape = self.shape if isinstance(self.shape, tuple) else (self.shape,)
        x = x.reshape((-1,) + shape)
        Sx = self.scattering(x)
        return Sx.reshape(batch_shape + (-1,))

    transform = predict
