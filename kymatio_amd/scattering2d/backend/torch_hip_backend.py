"""2D HIP backend namespace (plugin name 'torch'): the reference's op names
(kymatio/scattering2d/backend/torch_backend.py:19-183).  The frontend runs
the fused pipeline for this backend."""
from ...backend.torch_backend import TorchBackend


class TorchBackend2D(TorchBackend):
    @staticmethod
    def stack(arrays):
        return TorchBackend.stack(arrays, -3)


backend = TorchBackend2D  # alias module: backend="torch_hip"
