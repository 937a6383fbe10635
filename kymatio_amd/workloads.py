"""Seeded synthetic inputs for the scattering workloads (NumPy, CPU).

Every generator takes a ``numpy.random.Generator`` seed and returns float32
arrays, so the GPU run, the CPU baseline and the fixtures see the same data.
"""
import numpy as np


def uniform_images(n, side, seed=0):
    """n images side x side, pixels i.i.d. uniform[0, 1)."""
    rng = np.random.default_rng(seed)
    return rng.random((n, side, side), dtype=np.float32)


def grating_images(n, side, n_gratings=3, seed=0):
    """White noise N(0,1) plus ``n_gratings`` sinusoidal gratings per image.

    Each grating has a random orientation theta ~ U[0, pi), frequency
    f ~ U[0.02, 0.4] cycles/pixel, phase ~ U[0, 2 pi) and amplitude 1.
    """
    rng = np.random.default_rng(seed)
    v, u = np.meshgrid(np.arange(side, dtype=np.float64), np.arange(side, dtype=np.float64),
                       indexing='ij')
    out = rng.standard_normal((n, side, side))
    for i in range(n):
        for _ in range(n_gratings):
            f = rng.uniform(0.02, 0.4)
            th = rng.uniform(0.0, np.pi)
            ph = rng.uniform(0.0, 2 * np.pi)
            out[i] += np.sin(2 * np.pi * f * (u * np.cos(th) + v * np.sin(th)) + ph)
    return out.astype(np.float32)


def normal_images(n, side, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, side, side)).astype(np.float32)


def chirp_signals(n, length, seed=0, noise=0.05):
    """Sum of 1-8 linear chirps plus white noise of std ``noise``.

    Per chirp: start frequency f0 ~ U[0.01, 0.4] cycles/sample, slope
    s ~ U[-1e-4, 1e-4] cycles/sample^2, amplitude log-uniform in [0.1, 1],
    phase ~ U[0, 2 pi):  A cos(2 pi (f0 t + s t^2 / 2) + phase).
    """
    rng = np.random.default_rng(seed)
    t = np.arange(length, dtype=np.float64)
    out = noise * rng.standard_normal((n, length))
    for i in range(n):
        for _ in range(int(rng.integers(1, 9))):
            f0 = rng.uniform(0.01, 0.4)
            s = rng.uniform(-1e-4, 1e-4)
            a = np.exp(rng.uniform(np.log(0.1), 0.0))
            ph = rng.uniform(0, 2 * np.pi)
            out[i] += a * np.cos(2 * np.pi * (f0 * t + 0.5 * s * t * t) + ph)
    return out.astype(np.float32)


def mixed_size_images(n, sides=(64, 128, 256, 512), seed=0):
    """n square images with side drawn uniformly from ``sides``; pixels U[0,1)."""
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(sides), size=n)
    return [rng.random((sides[k], sides[k]), dtype=np.float32) for k in pick]
