"""Alias of the default 2D HIP backend (backend='torch_hip')."""
from .torch_backend import TorchBackend2D

backend = TorchBackend2D
