"""The GPU workgroup bodies, run single-threaded on the CPU by the host-only
harness libkymatio_emu.so, against the golden vectors and the oracle.  This
checks the plan, the digit-reversal maps, the twiddles and every index
formula of the 2D kernels (both the all-in-LDS route and, with a tiny LDS
cap, the column/row multi-pass route) without a GPU."""
import numpy as np
import pytest
import torch

from kymatio_amd.scattering2d.filter_bank import filter_bank
from kymatio_amd.scattering2d.utils import compute_padding
from oracle import scat2d
from tests.helpers import RTOL_CHANNEL, Emu, channel_rel_err, golden, golden_2d_cases, pack_filters_2d


@pytest.fixture(scope='module')
def emu():
    return Emu()


def _filters(shape, J, L, dt=np.float32):
    Mp, Np = compute_padding(shape[0], shape[1], J)
    fb = filter_bank(Mp, Np, J, L)
    return fb, pack_filters_2d(fb, dt)


def test_fixture_2d(emu):
    d = golden('test_data_2d.npz')
    x = d['x'].reshape(-1, 32, 32)
    fb, f = _filters((32, 32), 4, 8)
    for lds in (160 * 1024, 4096):
        out, info = emu.scat2d(x, 4, (32, 32), f, lds_limit=lds, chunk=2)
        assert np.allclose(out, d['Sx'].reshape(out.shape))
        if lds == 4096:
            assert info['big_nodes'] > 0


@pytest.mark.parametrize('case', golden_2d_cases(), ids=lambda c: c[0])
@pytest.mark.parametrize('lds', [160 * 1024, 8192], ids=['lds', 'multipass'])
def test_golden(emu, case, lds):
    tag, x, J, shape, kw, ref = case
    if tag == 'cfg2' and lds == 8192:
        pytest.skip('multi-pass route at 256x256 is covered on the GPU; too slow single-threaded')
    fb, f = _filters(shape, J, kw.get('L', 8))
    out, _ = emu.scat2d(x, J, shape, f, lds_limit=lds, **kw)
    assert out.shape == ref.shape
    assert channel_rel_err(out, ref) <= RTOL_CHANNEL


def test_fp64_matches_numpy_fp64(emu):
    G = golden('golden_scat2d.npz')
    x = G['cfg4_x']
    fb, f = _filters((128, 128), 4, 8, np.float64)
    out, _ = emu.scat2d(x, 4, (128, 128), f)
    ref = scat2d.scattering2d_numpy(x, 4, filters=fb)
    assert np.abs(out - ref).max() < 1e-13


def test_linearity_of_root_and_batch_independence(emu):
    """Images are independent: a batch equals its images run one at a time,
    and chunking does not change results."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5, 20, 24)).astype(np.float32)
    fb, f = _filters((20, 24), 2, 8)
    full, _ = emu.scat2d(x, 2, (20, 24), f)
    chunked, _ = emu.scat2d(x, 2, (20, 24), f, chunk=2)
    single = np.concatenate([emu.scat2d(x[i:i + 1], 2, (20, 24), f)[0] for i in range(5)])
    assert np.array_equal(full, chunked)
    assert np.array_equal(full, single)


def test_unsupported_prime_raises(emu):
    # M = 500, J = 2 -> P = 508 = 4 * 127: prime radix 127 is not built
    fb_dummy = np.zeros(1, np.float32)
    with pytest.raises(RuntimeError, match='prime factor'):
        emu.scat2d(np.zeros((1, 500, 500), np.float32), 2, (500, 500), fb_dummy)


def test_small_j_error(emu):
    with pytest.raises(RuntimeError, match='smallest dimension'):
        emu.scat2d(np.zeros((1, 10, 10), np.float32), 4, (10, 10), np.zeros(1, np.float32))


@pytest.mark.parametrize('shape,J,kw,lds', [
    ((20, 24), 2, {}, 160 * 1024),
    ((20, 24), 2, {}, 4096),
    ((16, 16), 2, {'pre_pad': 1}, 160 * 1024),
    ((31, 33), 3, {'max_order': 1}, 160 * 1024),
    ((32, 32), 4, {}, 8192),
    ((32, 32), 5, {}, 160 * 1024),
    ((8, 8), 2, {'L': 4}, 160 * 1024),
], ids=['lds', 'multipass', 'prepad', 'order1', 'J4-multipass', 'J5-fold', 'L4'])
def test_backward_matches_autograd_through_oracle(emu, shape, J, kw, lds):
    """FP64: the hand-written adjoint equals torch autograd through the
    op-for-op restatement of the reference (what the reference's own
    gradients are)."""
    L = kw.get('L', 8)
    fb, f = _filters(shape, J, L, np.float64)
    rng = np.random.default_rng(0)
    from kymatio_amd.scattering2d.utils import compute_padding
    xs = compute_padding(shape[0], shape[1], J) if kw.get('pre_pad') else shape
    x = rng.standard_normal((2,) + tuple(xs))
    sc = scat2d.Scattering2DTorchCPU(J, shape, L=L, max_order=kw.get('max_order', 2),
                                     pre_pad=bool(kw.get('pre_pad')), dtype=torch.float64, filters=fb)
    xt = torch.from_numpy(x).requires_grad_()
    S = sc(xt)
    w = rng.standard_normal(S.shape)
    (S * torch.from_numpy(w)).sum().backward()
    from tests.helpers import emu_train2d
    # three backward passes on one forward: the backward must not consume the saved state
    out, gx = emu_train2d(emu, x, J, shape, f, w, L=L, max_order=kw.get('max_order', 2),
                          pre_pad=kw.get('pre_pad', 0), lds_limit=lds, nbwd=3)
    g = xt.grad.numpy()
    assert np.abs(out - S.detach().numpy()).max() < 1e-13
    assert np.abs(gx - g).max() <= 1e-12 * np.abs(g).max()


@pytest.mark.parametrize('shape,J,tune', [((32, 32), 2, None), ((64, 64), 3, None), ((128, 128), 4, None),
                                          ((128, 128), 4, (16, 0, 16, 8)), ((128, 128), 4, (16, 0, 32, 8))])
def test_backward_fp32_compiled_nodes(emu, shape, J, tune):
    """FP32 backward through the compile-time node kernels (40², 20² and 80²
    nodes with 10² leaves) and, at 128² J=4, the compile-time 160² backward
    passes (the default blocks, and 16 or 32 columns by 8 rows): equals torch
    autograd through the FP64 oracle to FP32 rounding, on two backward passes
    of one forward."""
    fb64, _ = _filters(shape, J, 8, np.float64)
    _, f32 = _filters(shape, J, 8, np.float32)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2,) + shape)
    _, info = emu.scat2d(x[:1].astype(np.float32), J, shape, f32, tune=list(tune) if tune else None)
    assert info['bwd_spec_launches'] >= (11 if J == 4 else 1)  # at J=4: 6 node launches + 5 of the 160² passes
    sc = scat2d.Scattering2DTorchCPU(J, shape, dtype=torch.float64, filters=fb64)
    xt = torch.from_numpy(x).requires_grad_()
    S = sc(xt)
    w = rng.standard_normal(S.shape)
    (S * torch.from_numpy(w)).sum().backward()
    from tests.helpers import emu_train2d
    out, gx = emu_train2d(emu, x.astype(np.float32), J, shape, f32, w.astype(np.float32), nbwd=2, tune=tune)
    g = xt.grad.numpy()
    assert np.abs(out - S.detach().numpy()).max() <= 1e-5 * np.abs(S.detach().numpy()).max()
    assert np.abs(gx - g).max() <= 1e-5 * np.abs(g).max()
