"""Mixed-size batches: one HIP graph replay over per-size plans
(kym_scat2d_graph_*), checked against the oracle, the per-size module and
itself across replays."""
import numpy as np
import pytest
import torch

from kymatio_amd import workloads
from kymatio_amd.scattering2d.mixed import MixedScattering2D, default_J
from kymatio_amd.torch import Scattering2D
from oracle import scat2d
from tests.helpers import RTOL_CHANNEL, channel_rel_err, golden

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda')
SIDES = (64, 128, 256, 512)


def _images(counts, seed=3):
    rng = np.random.default_rng(seed)
    return {s: torch.from_numpy(rng.random((n, s, s), dtype=np.float32)).to(DEV) for s, n in counts.items()}


@pytest.fixture(scope='module')
def per_size():
    """Single-size modules (J = log2(side) - 3) for the eager comparisons."""
    return {s: Scattering2D(default_J(s), (s, s)).to(DEV) for s in SIDES}


def test_mixed_matches_oracle_per_size():
    """Four images of every size of the mixed workload vs the oracle: sides
    64 and 128 run the CPU oracle here, sides 256 and 512 compare with its
    outputs stored in tests/golden/golden_mixed2d.npz."""
    G = golden('golden_mixed2d.npz')
    groups = _images({64: 4, 128: 4})
    for s in (256, 512):
        x = np.random.default_rng(int(G['s%d_seed' % s])).random((4, s, s), dtype=np.float32)
        groups[s] = torch.from_numpy(x).to(DEV)
    mx = MixedScattering2D(autotune=False)
    out = mx.forward_grouped(groups)
    for s in SIDES:
        if s in (256, 512):
            ref = G['s%d_Sx' % s]
        else:
            with torch.no_grad():
                ref = scat2d.Scattering2DTorchCPU(default_J(s), (s, s))(groups[s].cpu()).numpy()
        got = out[s].cpu().numpy()
        assert got.shape == ref.shape == (4, {64: 217, 128: 417, 256: 681, 512: 1009}[s], 8, 8)
        err = channel_rel_err(got, ref)
        print('side %d: per-channel rel err %.3e' % (s, err))
        assert err < RTOL_CHANNEL


@pytest.mark.parametrize('tune', [[16, 0, 8, 8], [16, 0, 8, 8, 0, 1]])
def test_compiled_passes_large_sizes(tune):
    """Sides 256 (J=5) and 512 (J=6) with 8 x 8 pass blocks: the 160 / 320 /
    640 big-node passes run their compile-time kernels; vs the oracle's
    outputs in golden_mixed2d.npz."""
    G = golden('golden_mixed2d.npz')
    for s in (256, 512):
        J = int(G['s%d_J' % s])
        x = np.random.default_rng(int(G['s%d_seed' % s])).random((4, s, s), dtype=np.float32)
        sc = Scattering2D(J, (s, s)).to(DEV)
        sc.tune = list(tune)
        got = sc(torch.from_numpy(x).to(DEV)).cpu().numpy()
        err = channel_rel_err(got, G['s%d_Sx' % s])
        print('side %d tune %s: per-channel rel err %.3e' % (s, tune, err))
        assert err < RTOL_CHANNEL


def test_graph_replays_match_eager_with_work_between(per_size):
    mx = MixedScattering2D(autotune=False)
    counts = {64: 5, 128: 3, 256: 2, 512: 1}
    groups = _images(counts, seed=4)
    eager = {}
    with torch.no_grad():
        for s in SIDES:
            eager[s] = per_size[s](groups[s])
    first = mx.forward_grouped(groups)
    for rep in range(3):
        # eager work and synchronisation between replays
        junk = torch.randn(1 << 20, device=DEV)
        (junk * 2).sum().item()
        with torch.no_grad():
            per_size[64](groups[64][:1])
        torch.cuda.synchronize()
        again = mx.forward_grouped(groups)
        for s in SIDES:
            assert torch.equal(again[s], first[s]), (rep, s)
    for s in SIDES:
        assert torch.equal(first[s], eager[s]), s


def test_list_interface_and_new_counts(per_size):
    mx = MixedScattering2D(autotune=False)
    rng = np.random.default_rng(5)
    sides = [64, 256, 64, 128, 512, 128, 64]
    imgs = [torch.from_numpy(rng.random((s, s), dtype=np.float32)).to(DEV) for s in sides]
    out = mx(imgs)
    assert [o.shape[0] for o in out] == [{64: 217, 128: 417, 256: 681, 512: 1009}[s] for s in sides]
    with torch.no_grad():
        for i in (0, 1, 4):
            ref = per_size[sides[i]](imgs[i][None])[0]
            assert torch.equal(out[i], ref)
    # fewer images of each size replay the same graph (no rebuild)
    builds = mx.builds
    out2 = mx(imgs[:3])
    assert mx.builds == builds
    assert torch.equal(out2[1], out[1]) and torch.equal(out2[0], out[0])
    with pytest.raises(RuntimeError):
        mx([torch.zeros(64, 32, device=DEV)])


def test_new_counts_replay_without_rebuild(per_size):
    """Batches whose per-size counts differ (within the graph's maxima, and
    with a size absent) replay one graph; each result equals the per-size
    modules' bit for bit.  A count above a maximum rebuilds once."""
    mx = MixedScattering2D(autotune=False)
    batches = [{64: 6, 128: 5, 256: 3, 512: 2}, {64: 2, 128: 7, 256: 1, 512: 3}, {64: 9, 256: 4},
               {64: 1, 128: 1, 256: 1, 512: 1}]
    for i, counts in enumerate(batches):
        groups = _images(counts, seed=20 + i)
        out = mx.forward_grouped(groups)
        assert mx.builds == 1, (i, mx.builds)
        assert set(out) == set(counts)
        with torch.no_grad():
            for s, x in groups.items():
                assert torch.equal(out[s], per_size[s](x)), (i, s)
    grown = {64: 40, 512: 1}
    groups = _images(grown, seed=30)
    out = mx.forward_grouped(groups)
    assert mx.builds == 2
    with torch.no_grad():
        assert torch.equal(out[64], per_size[64](groups[64]))


def test_autotuned_mixed_workload(per_size):
    """The mixed workload (256 images, sides drawn from {64,128,256,512})
    with autotuned plans agrees with the per-size modules."""
    imgs = workloads.mixed_size_images(256, seed=6)
    by = {}
    for v in imgs:
        by.setdefault(v.shape[0], []).append(v)
    groups = {s: torch.from_numpy(np.stack(v)).to(DEV) for s, v in by.items()}
    assert sum(x.shape[0] for x in groups.values()) == 256
    mx = MixedScattering2D(autotune=True)
    out = mx.forward_grouped(groups)
    assert set(mx.tuning) == set(groups)
    with torch.no_grad():
        for s, x in groups.items():
            ref = per_size[s](x)
            err = channel_rel_err(out[s].cpu().numpy(), ref.cpu().numpy())
            print('side %d (%d images, tuning %s): rel err vs default plan %.3e' % (s, x.shape[0], mx.tuning[s], err))
            assert err < RTOL_CHANNEL
