"""backend='torch_ops': the same HIP ops, but the frontend runs the
reference's per-op core loop over them (one kernel per op) instead of the
fused pipeline -- the plugin route of kymatio/frontend/base_frontend.py:32-50."""
from .torch_backend import TorchBackend2D


class TorchOpsBackend2D(TorchBackend2D):
    name = 'torch_ops'
    fused = False


backend = TorchOpsBackend2D
