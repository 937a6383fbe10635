"""Oracle (test infrastructure only): 1D filter bank and frontend setup, restated.

Follows kymatio/scattering1d/filter_bank.py and the setup half of
kymatio/scattering1d/frontend/base_frontend.py (FP64 NumPy throughout):
  adaptive_choice_P          filter_bank.py:6-48
  morlet_1d / gauss_1d       filter_bank.py:51-102
  compute_sigma_psi          filter_bank.py:105-139
  compute_temporal_support   filter_bank.py:142-196
  get_max_dyadic_subsampling filter_bank.py:199-233
  compute_xi_max             filter_bank.py:236-251
  anden_generator            filter_bank.py:254-311
  scattering_filter_factory  filter_bank.py:322-404 (np.sum level reduction)
  compute_border_indices     scattering1d/utils.py:4-39
  compute_padding            scattering1d/utils.py:41-69
  parse_T                    scattering1d/utils.py:71-108
  setup (N_pad, pads, ind)   scattering1d/frontend/base_frontend.py:37-98
"""
import math
import numpy as np
from scipy.fft import ifft


def adaptive_choice_P(sigma, eps=1e-7):
    return int(math.ceil(math.sqrt(-2 * sigma ** 2 * math.log(eps)) + 1))


def morlet_1d(N, xi, sigma):
    P = min(adaptive_choice_P(sigma), 5)
    freqs = np.arange((1 - P) * N, P * N, dtype=float) / float(N)
    freqs_low = np.fft.fftfreq(N) if P == 1 else freqs
    low = np.exp(-(freqs_low ** 2) / (2 * sigma ** 2))
    low = low.reshape(2 * P - 1, -1).mean(axis=0)
    if xi:
        gab = np.exp(-(freqs - xi) ** 2 / (2 * sigma ** 2))
        gab = gab.reshape(2 * P - 1, -1).mean(axis=0)
        kappa = gab[0] / low[0]
        filt = gab - kappa * low
    else:
        filt = low
    filt /= np.abs(ifft(filt)).sum()
    return filt


def gauss_1d(N, sigma):
    return morlet_1d(N, None, sigma)


def compute_sigma_psi(xi, Q, r=math.sqrt(0.5)):
    factor = 1. / math.pow(2, 1. / Q)
    return xi * (1 - factor) / (1 + factor) * (1. / math.sqrt(2 * math.log(1. / r)))


def compute_temporal_support(h_f, criterion_amplitude=1e-3):
    import warnings
    h = ifft(h_f, axis=1)
    half = h.shape[1] // 2
    resid = np.fliplr(np.cumsum(np.fliplr(np.abs(h)[:, :half]), axis=1))
    ok = np.max(resid, axis=0) <= criterion_amplitude
    if np.any(ok):
        return int(np.min(np.where(ok)[0]) + 1)
    warnings.warn('Signal support is too small to avoid border effects')
    return half


def get_max_dyadic_subsampling(xi, sigma, alpha, **unused):
    upper = min(abs(xi) + alpha * sigma, 0.5)
    return int(math.floor(-math.log2(upper)) - 1)


def compute_xi_max(Q):
    return max(1. / (1. + math.pow(2., 3. / Q)), 0.35)


def anden_generator(J, Q, sigma0, r_psi, **unused):
    xi = compute_xi_max(Q)
    sigma = compute_sigma_psi(xi, Q, r=r_psi)
    sigma_min = sigma0 / 2 ** J
    if sigma <= sigma_min:
        xi = sigma
    else:
        yield xi, sigma
        while sigma > sigma_min * math.pow(2, 1 / Q):
            xi /= math.pow(2, 1 / Q)
            sigma /= math.pow(2, 1 / Q)
            yield xi, sigma
    elbow = xi
    for _ in range(Q - 1):
        xi -= 1 / Q * elbow
        yield xi, sigma_min


def scattering_filter_factory(N, J, Q, T, kw, reduction=np.sum):
    max_j, prev_J = 0, 0
    log2_T = math.floor(math.log2(T))
    banks = []
    for Ql in Q:
        bank = []
        for xi, sigma in anden_generator(J, Ql, **kw):
            levels = [morlet_1d(N, xi, sigma)]
            j = get_max_dyadic_subsampling(xi, sigma, **kw)
            for level in range(1, min(prev_J, j, 1 + log2_T)):
                levels.append(reduction(levels[0].reshape(2 ** level, -1), axis=0))
            bank.append({'levels': levels, 'xi': xi, 'sigma': sigma, 'j': j})
            max_j = max(j, max_j)
        prev_J = max_j
        banks.append(bank)
    sigma_low = kw['sigma0'] / T
    phi_levels = [gauss_1d(N, sigma_low)]
    for level in range(1, max(prev_J, 1 + log2_T)):
        phi_levels.append(reduction(phi_levels[0].reshape(2 ** level, -1), axis=0))
    phi = {'levels': phi_levels, 'xi': 0, 'sigma': sigma_low, 'j': log2_T, 'N': N}
    return tuple([phi] + banks)


def compute_border_indices(log2_T, J, i0, i1):
    start, end = {0: i0}, {0: i1}
    for j in range(1, max(log2_T, J) + 1):
        start[j] = start[j - 1] // 2 + start[j - 1] % 2
        end[j] = end[j - 1] // 2 + end[j - 1] % 2
    return start, end


def compute_padding(N, N_input):
    if N < N_input:
        raise ValueError('Padding support should be larger than the original signal size!')
    to_add = N - N_input
    left = to_add // 2
    right = to_add - left
    if max(left, right) >= N_input:
        raise ValueError('Too large padding value, will lead to NaN errors')
    return left, right


def parse_T(T, J, N_input):
    if T is None:
        return 2 ** J, 'local'
    if T == 'global':
        return 2 ** J, 'global'
    if T > N_input:
        raise ValueError('T too large')
    if T == 0:
        return 2 ** J, False
    if T < 1:
        raise ValueError('T must be ==0 or >=1')
    return T, 'local'


def setup(J, N_input, Q=1, T=None, stride=None, oversampling=0):
    """Setup half of ScatteringBase1D.build / create_filters
    (kymatio/scattering1d/frontend/base_frontend.py:37-104, 239-299)."""
    Qt = (Q, 1) if isinstance(Q, int) else (tuple(Q) + (1,) if len(Q) == 1 else tuple(Q))
    T_val, average = parse_T(T, J, N_input)
    log2_T = int(math.floor(math.log2(T_val)))
    if stride is None:
        log2_stride = log2_T
    else:
        log2_stride = int(math.log2(stride))
    if oversampling > 0:
        log2_stride = log2_T - oversampling
    kw = {'alpha': 5., 'r_psi': math.sqrt(0.5), 'sigma0': 0.1}
    phi_f = gauss_1d(N_input, kw['sigma0'] / T_val)
    min_to_pad = 3 * compute_temporal_support(phi_f.reshape(1, -1), criterion_amplitude=1e-3)
    J_max_support = int(np.floor(np.log2(3 * N_input - 2)))
    J_pad = min(int(np.ceil(np.log2(N_input + 2 * min_to_pad))), J_max_support)
    N_pad = 2 ** J_pad
    pad_left, pad_right = compute_padding(N_pad, N_input)
    ind_start, ind_end = compute_border_indices(log2_T, J, pad_left, pad_left + N_input)
    phi, psi1, psi2 = scattering_filter_factory(N_pad, J, Qt, T_val, kw)
    return {'N_pad': N_pad, 'pad_left': pad_left, 'pad_right': pad_right,
            'ind_start': ind_start, 'ind_end': ind_end, 'phi': phi, 'psi1': psi1,
            'psi2': psi2, 'log2_T': log2_T, 'log2_stride': log2_stride,
            'average': average, 'T': T_val, 'Q': Qt, 'oversampling': oversampling}
