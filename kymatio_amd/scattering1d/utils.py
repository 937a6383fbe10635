"""Padding and border indices (kymatio/scattering1d/utils.py:4-108)."""


def compute_border_indices(log2_T, J, i0, i1):
    """Start/end of the unpadded signal after each dyadic subsampling (rounding up)."""
    start, end = {0: i0}, {0: i1}
    for j in range(1, max(log2_T, J) + 1):
        start[j] = start[j - 1] // 2 + start[j - 1] % 2
        end[j] = end[j - 1] // 2 + end[j - 1] % 2
    return start, end


def compute_padding(N, N_input):
    if N < N_input:
        raise ValueError('Padding support should be larger than the original' + 'signal size!')
    to_add = N - N_input
    left = to_add // 2
    right = to_add - left
    if max(left, right) >= N_input:
        raise ValueError('Too large padding value, will lead to NaN errors')
    return left, right


def parse_T(T, J, N_input, T_alias='T'):
    if T is None:
        return 2 ** J, 'local'
    if T == 'global':
        return 2 ** J, 'global'
    if T > N_input:
        raise ValueError("The support {} of the low-pass filter cannot exceed input length (got {} > {}). "
                         "For large averaging size, consider passing {}='global'.".format(T_alias, T, N_input,
                                                                                         T_alias))
    if T == 0:
        return 2 ** J, False
    if T < 1:
        raise ValueError("{} must be ==0 or >=1 (got {})".format(T_alias, T))
    return T, 'local'
