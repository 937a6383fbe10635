"""The decoder oracle pinned by brute force: a direct loop over the 'same'
convolution of kernel 9 (zero padding 4), nearest upsampling and bf16
round-to-nearest-even written from their definitions."""
import numpy as np
import torch

from oracle import decoder as odec


def _bf16_rne(a):
    """Round float32 values to bf16 (nearest, ties to even) by bit arithmetic."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def _brute_layer(x, w, b, up, relu, scale=None):
    x = np.asarray(x, dtype=np.float32)
    if scale is not None:
        x = x * np.asarray(scale, dtype=np.float32)[None, :, None]
    B, Cin, T0 = x.shape
    T = T0 * up
    xu = np.empty((B, Cin, T), dtype=np.float32)
    for s in range(T):
        xu[:, :, s] = x[:, :, s // up]
    xr = _bf16_rne(xu)
    wr = _bf16_rne(w)  # (Cout, 9, Cin)
    Cout = w.shape[0]
    y = np.zeros((B, Cout, T))
    for o in range(Cout):
        for t in range(T):
            acc = np.full(B, float(b[o]))
            for j in range(9):
                s = t + j - 4
                if 0 <= s < T:
                    acc += xr[:, :, s] @ wr[o, j, :]
            y[:, o, t] = acc
    return np.maximum(y, 0) if relu else y


def test_layer_matches_brute_force():
    g = torch.Generator().manual_seed(5)
    for up, relu, scaled in ((1, True, True), (4, True, False), (1, False, False)):
        x = torch.randn(2, 3, 7, generator=g)
        L = {'w': torch.randn(4, 9, 3, generator=g), 'b': torch.randn(4, generator=g), 'up': up, 'relu': relu}
        scale = torch.rand(3, generator=g) + 0.5 if scaled else None
        got = odec.layer(x, L, scale).numpy()
        ref = _brute_layer(x.numpy(), L['w'].numpy(), L['b'].numpy(), up, relu,
                           None if scale is None else scale.numpy())
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (up, relu, scaled)


def test_bf16_rounding_ties_to_even():
    # 1 + 2^-8 lies halfway between bf16 neighbours 1 and 1 + 2^-7: ties go to the even mantissa (1)
    x = torch.tensor([[[1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8]]], dtype=torch.float32)
    L = {'w': torch.zeros(1, 9, 1), 'b': torch.zeros(1), 'up': 1, 'relu': False}
    L['w'][0, 4, 0] = 1.0
    y = odec.layer(x, L).numpy()[0, 0]
    assert y[0] == 1.0 and y[1] == 1.0 + 2 ** -6


def test_forward_chains_layers():
    g = torch.Generator().manual_seed(6)
    layers = [{'w': torch.randn(3, 9, 2, generator=g), 'b': torch.randn(3, generator=g), 'up': 1, 'relu': True},
              {'w': torch.randn(1, 9, 3, generator=g), 'b': torch.randn(1, generator=g), 'up': 2, 'relu': False}]
    S = torch.randn(2, 2, 5, generator=g)
    scale = torch.tensor([0.5, 2.0])
    out, outs = odec.forward(S, layers, scale)
    h = _brute_layer(S.numpy(), layers[0]['w'].numpy(), layers[0]['b'].numpy(), 1, True, scale.numpy())
    y = _brute_layer(h.astype(np.float32), layers[1]['w'].numpy(), layers[1]['b'].numpy(), 2, False)
    assert np.allclose(out.numpy(), y[:, 0, :], rtol=1e-9, atol=1e-9)
    assert len(outs) == 2
