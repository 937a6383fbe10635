"""3D solid harmonic wavelets and Gaussians in Fourier
(kymatio/scattering3d/filter_bank.py:7-184, kymatio/scattering3d/utils.py:
44-95): the same grids, normalisations and float32 / complex64 rounding, so
the filters are the reference's."""
import numpy as np
from scipy.special import factorial

try:
    from scipy.special import sph_harm_y as _sph

    def _ylm(l, m, polar, azimuthal):
        return _sph(l, m, polar, azimuthal)
except ImportError:  # older scipy
    from scipy.special import sph_harm as _sph_old

    def _ylm(l, m, polar, azimuthal):
        return _sph_old(m, l, azimuthal, polar)


def _grid(M, N, O, fourier):
    """ifftshifted integer grid (3, M, N, O) in float32; in radians / sample
    in Fourier."""
    g = np.fft.ifftshift(np.mgrid[-M // 2:-M // 2 + M, -N // 2:-N // 2 + N, -O // 2:-O // 2 + O].astype('float32'),
                         axes=(1, 2, 3))
    if fourier:
        g[0] *= 2 * np.pi / M
        g[1] *= 2 * np.pi / N
        g[2] *= 2 * np.pi / O
    return g


def _sqrt(x):
    with np.errstate(invalid='ignore'):
        return np.sqrt(x)


def get_3d_angles(cartesian_grid):
    """(polar, azimuthal) of a (3, M, N, O) grid ordered (z, y, x)."""
    z, y, x = cartesian_grid
    return np.arctan2(z, _sqrt(x ** 2 + y ** 2)) + np.pi / 2, np.arctan2(y, x)


def double_factorial(i):
    return 1 if i < 1 else np.prod(np.arange(i, 0, -2))


def gaussian_3d(M, N, O, sigma, fourier=True):
    g = _grid(M, N, O, fourier)
    s = 1. / sigma if fourier else sigma
    out = np.exp(-0.5 * (g ** 2).sum(0) / s ** 2)
    if not fourier:
        out /= (2 * np.pi) ** 1.5 * s ** 3
    return out


def solid_harmonic_3d(M, N, O, sigma, l, fourier=True):
    """The 2l + 1 wavelets of order l (m = -l .. l), complex64 (2l+1, M, N, O)."""
    g = _grid(M, N, O, fourier)
    s = 1. / sigma if fourier else sigma
    r2 = (g ** 2).sum(0)
    gauss = np.exp(-0.5 * r2 / s ** 2).astype('complex64')
    if l == 0:
        return gauss.reshape((1, M, N, O)) if fourier else gauss.reshape((1, M, N, O)) / ((2 * np.pi) ** 1.5 * s ** 3)
    poly_gauss = _sqrt(r2 ** l) * gauss / s ** l
    polar, azimuthal = get_3d_angles(g)
    out = np.zeros((2 * l + 1, M, N, O), np.complex64)
    for i, m in enumerate(range(-l, l + 1)):
        out[i] = _ylm(l, m, polar, azimuthal) * poly_gauss
    if l % 2 == 0:
        norm = 1. / (2 * np.pi * np.sqrt(l + 0.5) * double_factorial(l + 1))
    else:
        norm = 1. / (2 ** (0.5 * (l + 3)) * np.sqrt(np.pi * (2 * l + 1)) * factorial((l + 1) / 2))
    if fourier:
        norm *= (2 * np.pi) ** 1.5 * (-1j) ** l
    else:
        norm /= s ** 3
    out *= norm
    return out


def solid_harmonic_filter_bank(M, N, O, J, L, sigma_0, fourier=True):
    """filters[l]: complex64 (J+1, 2l+1, M, N, O), scale sigma_0 2^j."""
    filters = []
    for l in range(L + 1):
        f = np.zeros((J + 1, 2 * l + 1, M, N, O), dtype='complex64')
        for j in range(J + 1):
            f[j] = solid_harmonic_3d(M, N, O, sigma_0 * 2 ** j, l, fourier=fourier)
        filters.append(f)
    return filters


def gaussian_filter_bank(M, N, O, J, sigma_0, fourier=True):
    g = np.zeros((J + 1, M, N, O), dtype='complex64')
    for j in range(J + 1):
        g[j] = gaussian_3d(M, N, O, sigma_0 * 2 ** j, fourier=fourier)
    return g


def generate_weighted_sum_of_gaussians(grid, positions, weights, sigma):
    """Sum of weighted 3D Gaussians (kymatio/scattering3d/utils.py:5-42; the
    first zero weight ends a signal's list)."""
    _, M, N, O = grid.shape
    out = np.zeros((positions.shape[0], M, N, O))
    for b in range(positions.shape[0]):
        for c, w in zip(positions[b], weights[b]):
            if w == 0:
                break
            out[b] = out[b] + w * np.exp(-0.5 * ((grid[0] - c[0]) ** 2 + (grid[1] - c[1]) ** 2 +
                                                 (grid[2] - c[2]) ** 2) / sigma ** 2)
    return out / ((2 * np.pi) ** 1.5 * sigma ** 3)
