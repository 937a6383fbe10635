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
import os

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


def _psi_filter(M, N, J, L, j, theta):
    sig = morlet_2d(M, N, 0.8 * 2 ** j, (int(L - L / 2 - 1) - theta) * np.pi / L, 3.0 / 4.0 * np.pi / 2 ** j, 4.0 / L)
    sig_f = np.real(fft2(sig))
    return {'levels': [periodize_filter_fft(sig_f, r) for r in range(min(j + 1, max(J - 1, 1)))], 'j': j,
            'theta': theta}


def _phi_filter(M, N, J):
    phi_f = np.real(fft2(gabor_2d(M, N, 0.8 * 2 ** (J - 1), 0, 0)))
    return {'levels': [periodize_filter_fft(phi_f, r) for r in range(J)], 'j': J}


_CACHE = {}
_CACHE_MAX = 8


def filter_bank(M, N, J, L=8):
    """psi: J*L dicts {'levels', 'j', 'theta'} in (j, theta) order, levels
    r < min(j+1, max(J-1, 1)); phi: {'levels' (r < J), 'j': J}.

    The filters are independent, so they are built on a thread pool (numpy
    releases the GIL in its array loops; each filter's arithmetic is
    unchanged), and the result is kept per (M, N, J, L) so that several
    modules of one configuration build it once.  Callers get fresh copies."""
    key = (M, N, J, L)
    fb = _CACHE.get(key)
    if fb is None:
        from concurrent.futures import ThreadPoolExecutor
        jobs = [(j, theta) for j in range(J) for theta in range(L)]
        workers = max(1, min(len(jobs) + 1, os.cpu_count() or 1, 16))
        with ThreadPoolExecutor(workers) as ex:
            fphi = ex.submit(_phi_filter, M, N, J)
            psi = list(ex.map(lambda jt: _psi_filter(M, N, J, L, *jt), jobs))
            phi = fphi.result()
        fb = {'phi': phi, 'psi': psi}
        if len(_CACHE) >= _CACHE_MAX:
            _CACHE.pop(next(iter(_CACHE)))
        _CACHE[key] = fb
    return {'phi': {'levels': [a.copy() for a in fb['phi']['levels']], 'j': fb['phi']['j']},
            'psi': [{'levels': [a.copy() for a in p['levels']], 'j': p['j'], 'theta': p['theta']}
                    for p in fb['psi']]}


def leaf_taps(phi_levels, desc):
    """Taps and band tables of the separable low-pass leaves
    (kym_scat2d_leaf_taps in include/kymatio_hip.h).

    phi_levels[r] is the Fourier-domain low-pass filter at resolution r (the
    buffers the spectral leaf multiplies with); its spatial form
    phi_r = Re(ifft2(phi_levels[r])) is separable, phi_r[a][b] = g[a] h[b]
    with g = phi_r[:, 0] / sqrt(phi_r[0, 0]) and h = phi_r[0, :] / sqrt(phi_r[0, 0]).
    desc = (taps_off, taps_len, M_padded, N_padded, then per r < J and axis
    (y, then x): off, lo, cnt, band_off, kmax).  Per axis of n samples the
    taps are c[i] = g[(lo + i) mod n], i < cnt, and the band table is
    band[k][d] = c[(k - kl d) mod n] (0 where that index is >= cnt),
    k < kmax, d < 16, kl = 2^(J - r).  Returns the taps_len reals to store at
    taps_off of the packed filter buffer (float64).  Raises RuntimeError if a
    filter is not separable or the plan's tap window drops more than 1e-6 of
    its mass, which does not happen for the filter bank of this module.
    """
    taps_off, taps_len = int(desc[0]), int(desc[1])
    out = np.zeros(taps_len, np.float64)
    if taps_len == 0:
        return out
    J = (len(desc) - 4) // 10
    for r in range(J):
        ph = np.asarray(phi_levels[r], np.float64)
        sp = np.real(np.fft.ifft2(ph))
        c00 = sp[0, 0]
        if not c00 > 0:
            raise RuntimeError('low-pass filter %d has no positive centre tap' % r)
        g = sp[:, 0] / np.sqrt(c00)
        h = sp[0, :] / np.sqrt(c00)
        resid = np.abs(sp - np.outer(g, h)).max() / np.abs(sp).max()
        if resid > 1e-5:
            raise RuntimeError('low-pass filter %d is not separable (residual %.2e)' % (r, resid))
        kl = 2 ** (J - r)
        for v, base in ((g, 4 + 10 * r), (h, 4 + 10 * r + 5)):
            off, lo, cnt, band_off, kmax = [int(t) for t in desc[base:base + 5]]
            n = v.shape[0]
            kept = v[(lo + np.arange(cnt)) % n]
            dropped = 1.0 - np.abs(kept).sum() / np.abs(v).sum()
            if dropped > 1e-6:
                raise RuntimeError('leaf taps of level %d drop %.2e of the filter' % (r, dropped))
            out[off - taps_off:off - taps_off + cnt] = kept
            t = (np.arange(kmax)[:, None] - kl * np.arange(16)[None, :]) % n
            band = np.where(t < cnt, kept[np.minimum(t, cnt - 1)], 0.0)
            out[band_off - taps_off:band_off - taps_off + 16 * kmax] = band.reshape(-1)
    return out
