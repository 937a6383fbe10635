"""Alias of the default 1D HIP backend (backend='torch_hip')."""
from .torch_backend import TorchBackend1D

backend = TorchBackend1D
