"""Learned warm start for reconstructing 1D signals from their scattering
coefficients (decoder trained and run with the library's HIP kernels)."""
from .decoder import ScatteringDecoder1D
from .recon import reconstruct

__all__ = ['ScatteringDecoder1D', 'reconstruct']
