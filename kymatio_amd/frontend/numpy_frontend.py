"""NumPy frontend over the HIP path (kymatio/frontend/numpy_frontend.py:1-28).

The reference's NumPy frontend runs its core loop on the CPU with the numpy
backend.  Here a NumPy frontend object keeps a torch module of the same
transform and stages each call through device memory: the array is copied
to the device (float32 stays float32, every other dtype runs in float64, the
dtype the reference's numpy FFT computes in), the fused HIP forward runs, and
the result comes back as NumPy arrays.  There is no CPU fallback: without a
GPU and libkymatio_hip.so the call raises.
"""
import numpy as np
import torch


class ScatteringNumPy:
    def __init__(self):
        self.frontend_name = 'numpy'

    def __call__(self, x):
        """This method is an alias for `scattering`."""
        return self.scattering(x)

    @staticmethod
    def input_checks(x):
        """kymatio/backend/numpy_backend.py:11-20."""
        if x is None:
            raise TypeError('The input should be not empty.')
        if not isinstance(x, np.ndarray):
            raise TypeError(f'The input should be a numpy array, got type {type(x)}')

    def _device(self):
        if self.device is not None:
            return torch.device(self.device)
        return torch.device('cuda', torch.cuda.current_device())

    def _module(self, dtype):
        """The torch module for this dtype, on the device (built once)."""
        key = (str(self._device()), dtype)
        m = self._torch_modules.get(key)
        if m is None:
            m = self._make_torch_module().to(device=self._device(), dtype=dtype)
            self._torch_modules[key] = m
        return m

    def _stage_in(self, x):
        dtype = torch.float32 if x.dtype == np.float32 else torch.float64
        xt = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32 if dtype == torch.float32 else np.float64))
        return xt.to(self._device()), dtype

    @staticmethod
    def _stage_out(S):
        if torch.is_tensor(S):
            return S.cpu().numpy()
        if isinstance(S, dict):
            return {k: ScatteringNumPy._stage_out(v) for k, v in S.items()}
        if isinstance(S, list):
            return [ScatteringNumPy._stage_out(v) for v in S]
        return S
