"""Oracle (test infrastructure only): 2D Morlet filter bank, restated.

Follows kymatio/scattering2d/filter_bank.py step by step, with the same
NumPy dtypes so that rounding matches:
  * gabor_2d  -- kymatio/scattering2d/filter_bank.py:132-175
    complex64 accumulator, float32 rotation matrices, float64 curvature,
    5x5 periodic image sum, the constant 3.1415 in the normalisation.
  * morlet_2d -- :94-129   (wavelet minus K times the envelope, K = ratio of sums)
  * periodize_filter_fft -- :56-91  (mask, then the explicit alias-sum loop)
  * filter_bank -- :5-53   (psi for j<J, theta<L; levels r < min(j+1, max(J-1,1));
    phi = Gabor with sigma 0.8*2**(J-1), levels r < J)
"""
import numpy as np
from scipy.fft import fft2


def gabor_2d(M, N, sigma, theta, xi, slant=1.0, offset=0):
    """kymatio/scattering2d/filter_bank.py:132-175."""
    gab = np.zeros((M, N), np.complex64)
    c, s = np.cos(theta), np.sin(theta)
    rot = np.array([[c, -s], [s, c]], np.float32)
    rot_inv = np.array([[c, s], [-s, c]], np.float32)
    diag = np.array([[1, 0], [0, slant * slant]])
    curv = np.dot(rot, np.dot(diag, rot_inv)) / (2 * sigma * sigma)
    for ex in (-2, -1, 0, 1, 2):
        for ey in (-2, -1, 0, 1, 2):
            xx, yy = np.mgrid[offset + ex * M:offset + M + ex * M,
                              offset + ey * N:offset + N + ey * N]
            quad = (curv[0, 0] * np.multiply(xx, xx)
                    + (curv[0, 1] + curv[1, 0]) * np.multiply(xx, yy)
                    + curv[1, 1] * np.multiply(yy, yy))
            phase = xx * xi * np.cos(theta) + yy * xi * np.sin(theta)
            gab += np.exp(-quad + 1.j * phase)
    gab /= (2 * 3.1415 * sigma * sigma / slant)
    return gab


def morlet_2d(M, N, sigma, theta, xi, slant=0.5, offset=0):
    """kymatio/scattering2d/filter_bank.py:94-129."""
    wv = gabor_2d(M, N, sigma, theta, xi, slant, offset)
    env = gabor_2d(M, N, sigma, theta, 0, slant, offset)
    K = np.sum(wv) / np.sum(env)
    return wv - K * env


def periodize_filter_fft(x, res):
    """kymatio/scattering2d/filter_bank.py:56-91, loop for loop."""
    M, N = x.shape
    crop = np.zeros((M // 2 ** res, N // 2 ** res), x.dtype)
    mask = np.ones(x.shape, np.float32)
    len_x = int(M * (1 - 2 ** (-res)))
    start_x = int(M * 2 ** (-res - 1))
    len_y = int(N * (1 - 2 ** (-res)))
    start_y = int(N * 2 ** (-res - 1))
    mask[start_x:start_x + len_x, :] = 0
    mask[:, start_y:start_y + len_y] = 0
    x = np.multiply(x, mask)
    h, w = int(M / 2 ** res), int(N / 2 ** res)
    for k in range(h):
        for l in range(w):
            for i in range(int(2 ** res)):
                for j in range(int(2 ** res)):
                    crop[k, l] += x[k + i * h, l + j * w]
    return crop


def filter_bank(M, N, J, L=8):
    """kymatio/scattering2d/filter_bank.py:5-53."""
    psi = []
    for j in range(J):
        for theta in range(L):
            sig = morlet_2d(M, N, 0.8 * 2 ** j,
                            (int(L - L / 2 - 1) - theta) * np.pi / L,
                            3.0 / 4.0 * np.pi / 2 ** j, 4.0 / L)
            sig_f = np.real(fft2(sig))
            levels = [periodize_filter_fft(sig_f, r)
                      for r in range(min(j + 1, max(J - 1, 1)))]
            psi.append({'levels': levels, 'j': j, 'theta': theta})
    phi_sig = gabor_2d(M, N, 0.8 * 2 ** (J - 1), 0, 0)
    phi_f = np.real(fft2(phi_sig))
    phi = {'levels': [periodize_filter_fft(phi_f, r) for r in range(J)], 'j': J}
    return {'phi': phi, 'psi': psi}


def compute_padding(M, N, J):
    """kymatio/scattering2d/utils.py:4-24."""
    return (((M + 2 ** J) // 2 ** J + 1) * 2 ** J,
            ((N + 2 ** J) // 2 ** J + 1) * 2 ** J)
