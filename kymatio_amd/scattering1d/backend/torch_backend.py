"""1D HIP backend namespace (plugin name 'torch'); the frontend runs the fused
pipeline for it.  Op names of kymatio/scattering1d/backend/torch_backend.py."""
from ...backend.torch_backend import TorchBackend


class TorchBackend1D(TorchBackend):
    pass


backend = TorchBackend1D
