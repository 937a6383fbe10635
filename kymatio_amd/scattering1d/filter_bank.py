"""Morlet filter bank for Scattering1D, built on the CPU at construction.

Same construction as kymatio/scattering1d/filter_bank.py (FP64 NumPy):
morlet_1d/gauss_1d (:51-102) periodise a Gaussian over 2P-1 periods of the
frequency axis with P = min(ceil(sqrt(-2 sigma^2 log 1e-7) + 1), 5) (:6-48),
subtract kappa times the low-pass so the band-pass has zero mean, and
normalise by the l1 norm of the impulse response; centre frequencies and
bandwidths follow the constant-Q / constant-bandwidth progression of
anden_generator (:254-311) with compute_sigma_psi (:105-139) and
compute_xi_max (:236-251); the maximal dyadic subsampling of each filter is
get_max_dyadic_subsampling (:199-233); lower resolutions sum the 2^level
aliases (np.sum, :383-400).
"""
import math
import warnings

import numpy as np
from scipy.fft import ifft


def _periods(sigma, eps=1e-7):
    return int(math.ceil(math.sqrt(-2 * sigma ** 2 * math.log(eps)) + 1))


def morlet_1d(N, xi, sigma):
    P = min(_periods(sigma), 5)
    freqs = np.arange((1 - P) * N, P * N, dtype=float) / float(N)
    base = np.fft.fftfreq(N) if P == 1 else freqs
    low = np.exp(-(base ** 2) / (2 * sigma ** 2)).reshape(2 * P - 1, -1).mean(axis=0)
    if xi:
        band = np.exp(-(freqs - xi) ** 2 / (2 * sigma ** 2)).reshape(2 * P - 1, -1).mean(axis=0)
        out = band - (band[0] / low[0]) * low
    else:
        out = low
    out /= np.abs(ifft(out)).sum()
    return out


def gauss_1d(N, sigma):
    return morlet_1d(N, None, sigma)


def compute_sigma_psi(xi, Q, r=math.sqrt(0.5)):
    f = 1. / math.pow(2, 1. / Q)
    return xi * (1 - f) / (1 + f) * (1. / math.sqrt(2 * math.log(1. / r)))


def compute_temporal_support(h_f, criterion_amplitude=1e-3):
    h = ifft(h_f, axis=1)
    half = h.shape[1] // 2
    tail = np.fliplr(np.cumsum(np.fliplr(np.abs(h)[:, :half]), axis=1))
    below = np.max(tail, axis=0) <= criterion_amplitude
    if np.any(below):
        return int(np.min(np.where(below)[0]) + 1)
    warnings.warn('Signal support is too small to avoid border effects')
    return half


def get_max_dyadic_subsampling(xi, sigma, alpha, **unused):
    return int(math.floor(-math.log2(min(abs(xi) + alpha * sigma, 0.5))) - 1)


def compute_xi_max(Q):
    return max(1. / (1. + math.pow(2., 3. / Q)), 0.35)


def anden_generator(J, Q, sigma0, r_psi, **unused):
    xi = compute_xi_max(Q)
    sigma = compute_sigma_psi(xi, Q, r=r_psi)
    sigma_min = sigma0 / 2 ** J
    step = math.pow(2, 1 / Q)
    if sigma <= sigma_min:
        xi = sigma
    else:
        yield xi, sigma
        while sigma > sigma_min * step:
            xi /= step
            sigma /= step
            yield xi, sigma
    elbow = xi
    for _ in range(Q - 1):
        xi -= 1 / Q * elbow
        yield xi, sigma_min


def spin(filterbank_fn, filterbank_kwargs):
    """Filter bank with both signs of every center frequency
    (kymatio/scattering1d/filter_bank.py:314-319): the generator's (xi, sigma)
    pairs, then the same pairs with -xi; for frequency scattering, where the
    sign of xi is the direction (spin) of the wavelet along log-frequency."""
    def spinned(J, Q, **kwargs):
        pairs = list(filterbank_fn(J, Q, **kwargs))
        yield from pairs
        for xi, sigma in pairs:
            yield -xi, sigma
    return spinned, filterbank_kwargs


def scattering_filter_factory(N, J, Q, T, filterbank, _reduction=np.sum):
    fn, kw = filterbank
    max_j, prev_J = 0, 0
    log2_T = math.floor(math.log2(T))
    banks = []
    for Q_layer in Q:
        bank = []
        for xi, sigma in fn(J, Q_layer, **kw):
            levels = [morlet_1d(N, xi, sigma)]
            j = get_max_dyadic_subsampling(xi, sigma, **kw)
            for level in range(1, min(prev_J, j, 1 + log2_T)):
                levels.append(_reduction(levels[0].reshape(2 ** level, -1), axis=0))
            bank.append({'levels': levels, 'xi': xi, 'sigma': sigma, 'j': j})
            max_j = max(j, max_j)
        prev_J = max_j
        banks.append(bank)
    sigma_low = kw['sigma0'] / T
    phi_levels = [gauss_1d(N, sigma_low)]
    for level in range(1, max(prev_J, 1 + log2_T)):
        phi_levels.append(_reduction(phi_levels[0].reshape(2 ** level, -1), axis=0))
    return tuple([{'levels': phi_levels, 'xi': 0, 'sigma': sigma_low, 'j': log2_T, 'N': N}] + banks)
