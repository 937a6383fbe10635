"""Padding arithmetic (kymatio/scattering2d/utils.py:4-24,
kymatio/scattering2d/frontend/base_frontend.py:19-32)."""


def compute_padding(M, N, J):
    """Padded size: the next multiple of 2^J strictly above M + 2^J - (M mod 2^J)."""
    return (((M + 2 ** J) // 2 ** J + 1) * 2 ** J,
            ((N + 2 ** J) // 2 ** J + 1) * 2 ** J)


def pad_amounts(M, N, J):
    Mp, Np = compute_padding(M, N, J)
    return [(Mp - M) // 2, (Mp - M + 1) // 2, (Np - N) // 2, (Np - N + 1) // 2]
