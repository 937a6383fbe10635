"""Scattering2D on the GPU (through the C ABI) against the reference's
fixture, the golden vectors written from the reference and the oracle.

Parity bar (FP32): per output channel, max |ours - ref| <= 1e-5 * max |ref|
of that channel; the reference fixture is also checked with the reference's
own torch.allclose defaults.  FP64: against the NumPy FP64 oracle, 1e-12."""
import numpy as np
import pytest
import torch

import kymatio_amd
from kymatio_amd import workloads
from kymatio_amd.torch import Scattering2D
from oracle import scat2d
from tests.helpers import RTOL_CHANNEL, channel_rel_err, golden, golden_2d_cases

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def test_fixture_2d():
    """tests/scattering2d/test_torch_scattering2d.py:45-77 on the HIP path."""
    d = golden('test_data_2d.npz')
    x = torch.from_numpy(d['x']).to(DEV)
    S = torch.from_numpy(d['Sx']).to(DEV)
    J = int(d['J'])
    sc = kymatio_amd.Scattering2D(J, shape=(32, 32), pre_pad=bool(d['pre_pad']), frontend='torch').to(DEV)
    Sg = sc(x)
    assert Sg.shape == S.shape
    assert torch.allclose(Sg, S)
    sc1 = Scattering2D(J, shape=(32, 32), max_order=1).to(DEV)
    S1 = sc1(x)
    assert torch.allclose(S1, S[..., :S1.shape[-3], :, :])


@pytest.mark.parametrize('case', golden_2d_cases(), ids=lambda c: c[0])
@pytest.mark.parametrize('lds', [0, 8192], ids=['lds', 'multipass'])
def test_golden(case, lds):
    tag, x, J, shape, kw, ref = case
    print('case', tag, 'lds_limit', lds, flush=True)
    sc = Scattering2D(J, shape, **kw).to(DEV)
    sc.lds_limit = lds
    out = sc(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert out.shape == ref.shape
    err = channel_rel_err(out, ref)
    print('  per-channel rel err %.3e' % err)
    assert err <= RTOL_CHANNEL


def test_headline_error_by_order_vs_fp64():
    """256x256, J=3: error against the reference (golden) and, per order,
    against the FP64 NumPy restatement."""
    G = golden('golden_scat2d.npz')
    x = G['cfg2_x']
    sc = Scattering2D(3, (256, 256)).to(DEV)
    out = sc(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert channel_rel_err(out, G['cfg2_Sx']) <= RTOL_CHANNEL
    phi, psi = sc.load_filters()
    fb = {'phi': {'levels': [v.cpu().numpy()[..., 0] for v in phi['levels']]},
          'psi': [{'levels': [v.cpu().numpy()[..., 0] for v in p['levels']], 'j': p['j']} for p in psi]}
    ref64 = scat2d.scattering2d_numpy(x, 3, filters=fb)
    orders = scat2d.channel_orders(3, 8)
    for o in (0, 1, 2):
        e = channel_rel_err(out[:, orders == o], ref64[:, orders == o])
        print('order %d: per-channel rel err vs FP64 %.3e' % (o, e))
        assert e <= RTOL_CHANNEL


def test_headline_more_reference_images():
    """Three more reference-written images of the 256x256 J=3 workload
    (tests/golden/make_headline_golden.py), in one batch with the golden
    cfg2 image: per-channel bar, and the fraction of elements within the
    reference's own torch.allclose(rtol=1e-5, atol=1e-8)."""
    G = golden('golden_headline2d.npz')
    C = golden('golden_scat2d.npz')
    x = np.concatenate([C['cfg2_x'], G['x']])
    ref = np.concatenate([C['cfg2_Sx'], G['Sx']])
    sc = Scattering2D(3, (256, 256)).to(DEV)
    out = sc(torch.from_numpy(x).to(DEV)).cpu().numpy()
    err = channel_rel_err(out, ref)
    close = np.isclose(out, ref, rtol=1e-5, atol=1e-8)
    print('4 reference images: per-channel rel err %.3e, elements within allclose(1e-5, 1e-8): %.6f'
          % (err, close.mean()))
    assert err <= RTOL_CHANNEL
    assert close.mean() >= 0.99


def test_fp64_gpu_vs_numpy():
    G = golden('golden_scat2d.npz')
    x = G['cfg4_x'].astype(np.float64)
    sc = Scattering2D(4, (128, 128)).double().to(DEV)
    out = sc(torch.from_numpy(x).to(DEV)).cpu().numpy()
    phi, psi = sc.load_filters()
    fb = {'phi': {'levels': [v.cpu().numpy()[..., 0] for v in phi['levels']]},
          'psi': [{'levels': [v.cpu().numpy()[..., 0] for v in p['levels']], 'j': p['j']} for p in psi]}
    ref = scat2d.scattering2d_numpy(x, 4, filters=fb)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def test_full_batch_properties():
    """Config-2 size (64 x 256 x 256): the batch result equals per-image runs
    bit for bit, is independent of the chunking, deterministic, and finite."""
    x = torch.from_numpy(workloads.grating_images(64, 256, seed=2)).to(DEV)
    sc = Scattering2D(3, (256, 256)).to(DEV)
    a = sc(x)
    b = sc(x)
    assert torch.equal(a, b)
    sc.images_per_chunk = 5
    c = sc(x)
    assert torch.equal(a, c)
    # chunks rotate over the caller's stream and the plan's streams (two by
    # default, up to four) or stay on the caller's stream: the same bits; the
    # plan's streams wait for the caller's pending work (the clone) and the
    # caller's stream for them (the result read right after)
    sc.images_per_chunk = 8
    d2 = sc(x.clone() * 1.0)
    sc.tune = [0, 0, 0, 0, 0, 1]
    d1 = sc(x)
    sc.tune = [0, 0, 0, 0, 0, 4]
    sc._plans.clear()
    d4 = sc(x * 1.0)
    sc.tune = [0, 0, 0, 0, 0, 3]
    sc._plans.clear()
    d3 = sc(x * 1.0)
    assert torch.equal(a, d1) and torch.equal(a, d2) and torch.equal(a, d3) and torch.equal(a, d4)
    sc._plans.clear()
    sc.tune = None
    side = torch.cuda.Stream(device=DEV)
    side.wait_stream(torch.cuda.current_stream(DEV))
    with torch.cuda.stream(side):
        e = sc(x + 0.0)
    torch.cuda.current_stream(DEV).wait_stream(side)
    assert torch.equal(a, e)
    sc.images_per_chunk = None
    for i in (0, 31, 63):
        assert torch.equal(sc(x[i:i + 1].contiguous())[0], a[i])
    assert torch.isfinite(a).all()
    assert a.shape == (64, 217, 32, 32)


def test_batch_shape_agnostic():
    """tests/scattering2d/test_torch_scattering2d.py:91-136."""
    S = Scattering2D(3, (32, 32), 8).to(DEV)
    with pytest.raises(RuntimeError, match='at least two'):
        S(torch.zeros((), device=DEV))
    with pytest.raises(RuntimeError, match='at least two'):
        S(torch.zeros((32,), device=DEV))
    Sx = S(torch.zeros((32, 32), device=DEV))
    assert Sx.shape == (217, 4, 4)
    for shp in ((1, 32, 32), (2, 32, 32), (2, 2, 32, 32), (2, 2, 2, 32, 32)):
        Sx = S(torch.zeros(shp, device=DEV))
        assert Sx.shape == shp[:-2] + (217, 4, 4)


def test_errors():
    """tests/scattering2d/test_torch_scattering2d.py:138-177."""
    S = Scattering2D(3, (32, 32)).to(DEV)
    with pytest.raises(TypeError, match='input should be'):
        S(None)
    x = torch.randn(4, 4, device=DEV)
    with pytest.raises(RuntimeError, match='must be contiguous'):
        S(x[::2, ::2])
    with pytest.raises(RuntimeError, match='Tensor must be of spatial size'):
        S(torch.randn(31, 31, device=DEV))
    S = Scattering2D(3, (32, 32), pre_pad=True).to(DEV)
    with pytest.raises(RuntimeError, match='Padded tensor must be of spatial size'):
        S(torch.randn(31, 31, device=DEV))
    S = Scattering2D(2, (8, 8)).to(DEV)
    y = S(torch.randn(8, 8, device=DEV))
    assert y.device.type == 'cuda'
    with pytest.raises(TypeError, match='same dtype'):
        S(torch.randn(8, 8, device=DEV, dtype=torch.float64))


def test_input_size_agnostic():
    """tests/scattering2d/test_torch_scattering2d.py:179-222 (odd sizes, J up to 5)."""
    for N in (31, 32, 33):
        for J in (1, 2, 4):
            sc = Scattering2D(J, shape=(N, N)).to(DEV)
            sc(torch.zeros(3, 3, N, N, device=DEV))
            sc = Scattering2D(J, shape=(N, N), pre_pad=True).to(DEV)
            sc(torch.zeros(3, 3, sc._M_padded, sc._N_padded, device=DEV))
    sc = Scattering2D(5, shape=(32, 32)).to(DEV)
    assert sc(torch.zeros(3, 3, 32, 32, device=DEV)).shape[-2:] == (1, 1)
    sc = Scattering2D(5, shape=(37, 32)).to(DEV)
    assert sc(torch.zeros(3, 3, 37, 32, device=DEV)).shape[-2:] == (1, 1)


def test_list_output_meta():
    sc = Scattering2D(2, (16, 16), out_type='list').to(DEV)
    x = torch.randn(2, 16, 16, device=DEV)
    out = sc(x)
    arr = Scattering2D(2, (16, 16)).to(DEV)(x)
    assert len(out) == arr.shape[-3]
    ref = scat2d.Scattering2DTorchCPU(2, (16, 16))(x.cpu(), out_type='list')
    for o, r, c in zip(out, ref, range(arr.shape[-3])):
        assert o['j'] == r['j'] and o['n'] == r['n'] and o['theta'] == r['theta']
        assert torch.equal(o['coef'], arr[:, c])
