"""Morlet filter bank for Scattering2D, built on the CPU at construction.

Same construction as kymatio/scattering2d/filter_bank.py (filter_bank :5-53,
periodize_filter_fft :56-91, morlet_2d :94-129, gabor_2d :132-175) and the
same NumPy dtypes at every step, so the float32 filters are bit-identical to
the reference's: complex64 accumulation of the 5 x 5 periodic images, float32
rotation matrices, the 3.1415 normalisation constant, the real part of the
complex64 FFT, and the alias sums of the crop accumulated in the same
(row alias, column alias) order.  The crop is vectorised over pixels instead
of the reference's per-pixel Python loop.
"""
import numpy as np
from scipy.fft import fft2


def gabor_2d(M, N, sigma, theta, xi, slant=1.0, offset=0):
    c, s = np.cos(theta), np.sin(theta)
    rot = np.array([[c, -s], [s, c]], np.float32)
    rot_inv = np.array([[c, s], [-s, c]], np.float32)
    curv = np.dot(rot, np.dot(np.array([[1, 0], [0, slant * slant]]), rot_inv)) / (2 * sigma * sigma)
    acc = np.zeros((M, N), np.complex64)
    rows = np.arange(M, dtype=np.int64)[:, None]
    cols = np.arange(N, dtype=np.int64)[None, :]
    for ex in (-2, -1, 0, 1, 2):
        xx = np.broadcast_to(offset + ex * M + rows, (M, N))
        for ey in (-2, -1, 0, 1, 2):
            yy = np.broadcast_to(offset + ey * N + cols, (M, N))
            arg = -(curv[0, 0] * np.multiply(xx, xx) + (curv[0, 1] + curv[1, 0]) * np.multiply(xx, yy)
                    + curv[1, 1] * np.multiply(yy, yy)) + 1.j * (xx * xi * np.cos(theta) + yy * xi * np.sin(theta))
            acc += np.exp(arg)
    acc /= (2 * 3.1415 * sigma * sigma / slant)
    return acc


def morlet_2d(M, N, sigma, theta, xi, slant=0.5, offset=0):
    wave = gabor_2d(M, N, sigma, theta, xi, slant, offset)
    envelope = gabor_2d(M, N, sigma, theta, 0, slant, offset)
    return wave - (np.sum(wave) / np.sum(envelope)) * envelope


def periodize_filter_fft(x, res):
    """Crop a Fourier-domain filter to resolution 2^-res by summing its aliases
    (after zeroing the band the crop cannot represent)."""
    M, N = x.shape
    h, w = int(M / 2 ** res), int(N / 2 ** res)
    mask = np.ones(x.shape, np.float32)
    sx, lx = int(M * 2 ** (-res - 1)), int(M * (1 - 2 ** (-res)))
    sy, ly = int(N * 2 ** (-res - 1)), int(N * (1 - 2 ** (-res)))
    mask[sx:sx + lx, :] = 0
    mask[:, sy:sy + ly] = 0
    x = np.multiply(x, mask)
    crop = np.zeros((M // 2 ** res, N // 2 ** res), x.dtype)
    for i in range(int(2 ** res)):
        for j in range(int(2 ** res)):
            crop[:h, :w] += x[i * h:(i + 1) * h, j * w:(j + 1) * w]
    return crop


def filter_bank(M, N, J, L=8):
    """psi: J*L dicts {'levels', 'j', 'theta'} in (j, theta) order, levels
    r < min(j+1, max(J-1, 1)); phi: {'levels' (r < J), 'j': J}."""
    psi = []
    for j in range(J):
        for theta in range(L):
            sig = morlet_2d(M, N, 0.8 * 2 ** j, (int(L - L / 2 - 1) - theta) * np.pi / L,
                            3.0 / 4.0 * np.pi / 2 ** j, 4.0 / L)
            sig_f = np.real(fft2(sig))
            psi.append({'levels': [periodize_filter_fft(sig_f, r) for r in range(min(j + 1, max(J - 1, 1)))],
                        'j': j, 'theta': theta})
    phi_f = np.real(fft2(gabor_2d(M, N, 0.8 * 2 ** (J - 1), 0, 0)))
    phi = {'levels': [periodize_filter_fft(phi_f, r) for r in range(J)], 'j': J}
    return {'phi': phi, 'psi': psi}
