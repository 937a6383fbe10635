"""Product filter bank (vectorised) against the oracle's literal restatement
and the reference's golden checksums."""
import numpy as np
import pytest

from kymatio_amd.scattering2d.filter_bank import filter_bank
from oracle import filters2d
from tests.helpers import golden


@pytest.mark.parametrize('M,N,J,L', [(40, 40, 2, 8), (64, 64, 4, 8), (34, 34, 1, 8), (48, 40, 3, 6)])
def test_bit_identical_to_oracle(M, N, J, L):
    a = filter_bank(M, N, J, L)
    b = filters2d.filter_bank(M, N, J, L)
    la = list(a['phi']['levels']) + [l for p in a['psi'] for l in p['levels']]
    lb = list(b['phi']['levels']) + [l for p in b['psi'] for l in p['levels']]
    assert len(la) == len(lb)
    for x, y in zip(la, lb):
        assert x.dtype == y.dtype == np.float32
        assert np.array_equal(x, y)


def test_P272_matches_reference_golden():
    g = golden('golden_filters2d.npz')
    fb = filter_bank(272, 272, 3, 8)
    allf = [fb['phi']['levels'][r] for r in range(3)] + [l for p in fb['psi'] for l in p['levels']]
    assert np.array_equal(np.array([a.astype(np.float64).sum() for a in allf]), g['P272_sum'])
    vals = np.stack([a.reshape(-1)[i] for a, i in zip(allf, g['P272_idx'])])
    assert np.array_equal(vals, g['P272_val'])
