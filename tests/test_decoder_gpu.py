"""Learned-decoder kernels on the GPU: the bf16 MFMA lane layouts, and the
decoder's forward and hand-written backward against the CPU oracle."""
import ctypes

import numpy as np
import pytest
import torch

from kymatio_amd import _lib
from kymatio_amd.learned import ScatteringDecoder1D
from oracle import decoder as odec

pytestmark = pytest.mark.gpu
DEV = 'cuda'


def test_mfma_lane_layouts():
    out16 = torch.zeros(256, device=DEV)
    out32 = torch.zeros(256, device=DEV)
    _lib.check(_lib.lib().kym_mfma_layout_check(ctypes.c_void_p(out16.data_ptr()), ctypes.c_void_p(out32.data_ptr()),
                                                 _lib.stream_of(out16.device)))
    m = np.arange(16)
    for K, out in ((16, out16), (32, out32)):
        k = np.arange(K)
        A = (7 * m[:, None] + 3 * k[None, :]) % 17
        Bm = (5 * k[:, None] + 11 * m[None, :]) % 13
        ref = (A @ Bm).astype(np.float32)
        got = out.cpu().numpy().reshape(16, 16)
        assert np.array_equal(got, ref), K


def _small_decoder():
    dec = ScatteringDecoder1D(n_channels=20, t_in=64, n_out=4096, width=64, device=DEV, seed=1)
    g = torch.Generator().manual_seed(2)
    for L in dec.layers:
        L['b'] = (0.1 * torch.randn(L['cout'], generator=g)).to(DEV)
    dec._pack()
    return dec


def _cpu_layers(dec):
    return [{k: (v.cpu() if torch.is_tensor(v) else v) for k, v in L.items()} for L in dec.layers]


def test_forward_layers_match_oracle():
    """Each layer from the same FP32 input: both sides round the same values
    to bf16 (round to nearest even), so only the FP32 accumulation order
    separates them."""
    dec = _small_decoder()
    S = torch.rand(3, 20, 64, device=DEV)
    dec.fit_scale(S)
    _, acts = dec.forward(S, keep=True)
    for i, L in enumerate(_cpu_layers(dec)):
        ref = odec.layer(acts[i].cpu(), L, dec.scale.cpu() if i == 0 else None)
        rel = (acts[i + 1].cpu().double() - ref).abs().max() / ref.abs().max()
        print('layer %d max rel err %.3e' % (i, rel))
        assert rel < 2e-5


def test_forward_end_to_end_matches_oracle():
    """Whole network: an FP32 sum that lands a hair either side of a bf16
    rounding boundary moves that activation by one bf16 step (2^-8
    relative), so the end-to-end bar is at bf16 resolution."""
    dec = _small_decoder()
    S = torch.rand(3, 20, 64, device=DEV)
    dec.fit_scale(S)
    y = dec.forward(S)
    ref, _ = odec.forward(S.cpu(), _cpu_layers(dec), dec.scale.cpu())
    rel = (y.cpu().double() - ref).abs().max() / ref.abs().max()
    print('forward max rel err %.3e' % rel)
    assert rel < 2e-2


def test_backward_matches_oracle():
    dec = _small_decoder()
    S = torch.rand(2, 20, 64, device=DEV)
    dec.fit_scale(S)
    y, acts = dec.forward(S, keep=True)
    dy = torch.randn_like(y)
    grads = dec.backward(dy, acts)
    layers = []
    for L in dec.layers:
        d = {k: (v.cpu() if torch.is_tensor(v) else v) for k, v in L.items()}
        d['w'] = d['w'].double().requires_grad_()
        d['b'] = d['b'].double().requires_grad_()
        layers.append(d)
    ref, _ = odec.forward(S.cpu(), layers, dec.scale.cpu())
    (ref * dy.cpu().double()).sum().backward()
    for i, (L, (dW, db)) in enumerate(zip(layers, grads)):
        rw = L['w'].grad
        ew = (dW.cpu().double() - rw).abs().max() / rw.abs().max()
        eb = (db.cpu().double() - L['b'].grad).abs().max() / L['b'].grad.abs().max()
        print('layer %d: dW rel err %.3e, db rel err %.3e' % (i, ew, eb))
        assert ew < 3e-2 and eb < 3e-2


def test_adam_reduces_loss_on_fixed_batch():
    dec = _small_decoder()
    S = torch.rand(4, 20, 64, device=DEV)
    dec.fit_scale(S)
    target = torch.randn(4, 4096, device=DEV) * 0.1
    losses = []
    for _ in range(30):
        y, acts = dec.forward(S, keep=True)
        loss = ((y - target) ** 2).mean()
        losses.append(loss.item())
        grads = dec.backward(2 * (y - target) / y.numel(), acts)
        dec.adam_step(grads, lr=1e-3)
    assert losses[-1] < 0.9 * losses[0]


def test_chirp_generator():
    x = torch.empty(4, 4096, device=DEV)
    _lib.check(_lib.lib().kym_chirps(ctypes.c_void_p(x.data_ptr()), 4, 4096, 7, 0, 0.05, _lib.stream_of(x.device)))
    y = torch.empty_like(x)
    _lib.check(_lib.lib().kym_chirps(ctypes.c_void_p(y.data_ptr()), 4, 4096, 7, 0, 0.05, _lib.stream_of(y.device)))
    assert torch.equal(x, y)
    assert torch.isfinite(x).all()
    assert 0.02 < x.std().item() < 3.0
