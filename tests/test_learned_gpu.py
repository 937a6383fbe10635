"""What the trained warm-start decoder gains (Scattering1D J=6 Q=8 N=4096,
chirp signals): its output starts far closer to the target scattering than
an untrained decoder or N(0, 1) noise, and Adam refinement through the HIP
backward from it reaches a given loss in fewer iterations than from noise
at the same step size."""
import os

import pytest
import torch

from kymatio_amd.learned import ScatteringDecoder1D, reconstruct
from kymatio_amd.learned.evaluate import load_decoder
from kymatio_amd.learned.recon import rel_error
from kymatio_amd.learned.train import CHECKPOINT, chirps
from kymatio_amd.torch import Scattering1D

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda')


@pytest.fixture(scope='module')
def setup():
    sc = Scattering1D(6, 4096, 8).to(DEV)
    x = chirps(64, 4096, 4321, 0, device=DEV)  # signals never seen in training (other seed)
    with torch.no_grad():
        S = sc(x)
    return sc, x, S


def test_checkpoint_loads_safely():
    assert os.path.exists(CHECKPOINT)
    state = torch.load(CHECKPOINT, map_location='cpu', weights_only=True)
    assert [int(v) for v in state['config']] == [126, 64, 4096, 64]
    dec = load_decoder(CHECKPOINT, DEV)
    assert 180_000 < dec.n_params() < 260_000


def test_trained_start_is_closer(setup):
    sc, x, S = setup
    dec = load_decoder(CHECKPOINT, DEV)
    with torch.no_grad():
        warm = dec.forward(S)
    untrained = ScatteringDecoder1D(126, 64, 4096, device=DEV, seed=7)
    untrained.fit_scale(S)
    with torch.no_grad():
        cold = untrained.forward(S)
    noise = torch.randn(x.shape, device=DEV, generator=torch.Generator(device=DEV).manual_seed(0))
    r_warm, r_cold, r_noise = rel_error(sc, warm, S), rel_error(sc, cold, S), rel_error(sc, noise, S)
    print('start ||S(y)-S(x)||/||S(x)||: trained %.3e, untrained %.3e, noise %.3e' % (r_warm, r_cold, r_noise))
    assert r_warm < 0.55
    assert r_warm < 0.55 * r_cold and r_warm < 0.45 * r_noise


def test_refinement_gain_at_equal_step(setup):
    sc, x, S = setup
    dec = load_decoder(CHECKPOINT, DEV)
    with torch.no_grad():
        warm = dec.forward(S)
    noise = torch.randn(x.shape, device=DEV, generator=torch.Generator(device=DEV).manual_seed(1))
    _, hw, _ = reconstruct(sc, S, warm, tol=0.0, max_iter=300, method='adam', lr=0.1, check_every=50)
    _, hn, _ = reconstruct(sc, S, noise, tol=0.0, max_iter=300, method='adam', lr=0.1, check_every=50)
    print('after 300 Adam steps (lr 0.1): from trained start %.3e, from noise %.3e' % (hw[-1][2], hn[-1][2]))
    assert hw[-1][2] < hn[-1][2]
