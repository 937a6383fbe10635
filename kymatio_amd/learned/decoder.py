"""1D convolutional decoder S(x) -> x on bf16 MFMA kernels (kym_dec_* in
include/kymatio_hip.h).

Architecture (about 221k parameters): the scattering coefficients
(C channels x T_S samples) pass through six 'same' convolutions of kernel 9,
64 channels, ReLU between layers and nearest upsampling by 4 before layers
2-4, ending in one output channel at the signal length:
    C@T_S -> 64@T_S -> (x4) 64@4T_S -> (x4) 64@16T_S -> (x4) 64@64T_S -> 64 -> 1
Weights are FP32 masters (layout (Cout, 9, Cin)) updated by Adam and packed
to bf16 for the MFMA forward / backward each step.  The input is scaled per
channel by 1 / rms of the training coefficients.
"""
import ctypes
import math

import torch

from .. import _lib

_P = ctypes.c_void_p


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _pad16(c):
    return (c + 15) // 16 * 16


def _pad64(c):
    return (c + 63) // 64 * 64


class ScatteringDecoder1D:
    def __init__(self, n_channels, t_in, n_out, width=64, device='cuda', seed=0):
        self.device = torch.device(device)
        self.n_channels, self.t_in, self.n_out, self.width = n_channels, t_in, n_out, width
        ups = [1, 4, 4, 4, 1, 1]
        T = t_in
        self.layers = []
        g = torch.Generator().manual_seed(seed)
        cin = n_channels
        for i, up in enumerate(ups):
            T *= up
            cout = 1 if i == len(ups) - 1 else width
            fan_in = cin * 9
            w = torch.randn(cout, 9, cin, generator=g) * math.sqrt(2.0 / fan_in)
            if i == len(ups) - 1:
                w *= 0.5
            self.layers.append({'cin': cin, 'cout': cout, 'T': T, 'up': up, 'relu': i < len(ups) - 1,
                                'w': w.to(self.device), 'b': torch.zeros(cout, device=self.device),
                                'cin_pad': _pad64(cin), 'cout_pad': _pad16(cout)})
            cin = cout
        if T != n_out:
            raise ValueError('decoder output length %d != signal length %d' % (T, n_out))
        self.scale = torch.ones(n_channels, device=self.device)
        self._adam = None
        self.step_count = 0
        self._pack()

    # ------------------------------------------------------------ parameters
    def parameters(self):
        for L in self.layers:
            yield L['w']
            yield L['b']

    def n_params(self):
        return sum(p.numel() for p in self.parameters())

    def state_dict(self):
        d = {'scale': self.scale.cpu()}
        for i, L in enumerate(self.layers):
            d['w%d' % i] = L['w'].cpu()
            d['b%d' % i] = L['b'].cpu()
        d['config'] = torch.tensor([self.n_channels, self.t_in, self.n_out, self.width])
        return d

    def load_state_dict(self, d):
        self.scale = d['scale'].to(self.device)
        for i, L in enumerate(self.layers):
            L['w'] = d['w%d' % i].to(self.device)
            L['b'] = d['b%d' % i].to(self.device)
        self._pack()

    def _pack(self):
        lib = _lib.lib()
        s = _lib.stream_of(self.device)
        for L in self.layers:
            L['wf'] = torch.empty(L['cout_pad'] * 9 * L['cin_pad'], dtype=torch.int16, device=self.device)
            L['wt'] = torch.zeros(L['cin_pad'] * 9 * L['cout_pad'], dtype=torch.int16, device=self.device)
            _lib.check(lib.kym_dec_pack_weights(_p(L['w']), _p(L['wf']), _p(L['wt']), L['cout'], L['cout_pad'],
                                                 L['cin'], L['cin_pad'], s))

    def fit_scale(self, S):
        """Per-channel input scale 1 / rms over a batch of coefficients (B, C, T)."""
        rms = S.float().pow(2).mean(dim=(0, 2)).sqrt()
        self.scale = (1.0 / rms.clamp_min(1e-12)).contiguous()

    # ------------------------------------------------------------ forward / backward
    def forward(self, S, keep=False):
        """S: (B, C, T_S) float32 on the device -> (B, n_out).  With keep=True
        also returns the activations needed by backward()."""
        lib = _lib.lib()
        s = _lib.stream_of(self.device)
        B = S.shape[0]
        x = S.contiguous()
        acts = [x]
        for i, L in enumerate(self.layers):
            y = torch.empty(B, L['cout'], L['T'], device=self.device)
            _lib.check(lib.kym_dec_conv_fwd(_p(x), _p(L['wf']), _p(L['b']), _p(self.scale) if i == 0 else None,
                                             _p(y), B, L['cin'], L['cin_pad'], L['cout'], L['cout_pad'], L['T'],
                                             L['up'], int(L['relu']), s))
            x = y
            acts.append(y)
        out = x.reshape(B, self.n_out)
        return (out, acts) if keep else out

    def backward(self, dy, acts):
        """dy: (B, n_out) gradient of the loss w.r.t. the output; returns
        [(dW, db)] per layer (FP32, master layout)."""
        lib = _lib.lib()
        s = _lib.stream_of(self.device)
        B = dy.shape[0]
        g = dy.reshape(B, 1, self.n_out).contiguous()
        grads = [None] * len(self.layers)
        for i in range(len(self.layers) - 1, -1, -1):
            L = self.layers[i]
            xin = acts[i]
            dW = torch.zeros(L['cout'], 9, L['cin'], device=self.device)
            db = torch.zeros(L['cout'], device=self.device)
            _lib.check(lib.kym_dec_conv_bwd_weight(_p(g), _p(xin), _p(self.scale) if i == 0 else None, _p(dW), B,
                                                    L['cin'], L['cin_pad'], L['cout'], L['cout_pad'], L['T'],
                                                    L['up'], 16, s))
            _lib.check(lib.kym_dec_bias_grad(_p(g), _p(db), B, L['cout'], L['T'], s))
            grads[i] = (dW, db)
            if i == 0:
                break
            # dX at the layer's (upsampled) resolution: transposed conv = conv with flipped, transposed W
            dx = torch.empty(B, L['cin'], L['T'], device=self.device)
            _lib.check(lib.kym_dec_conv_fwd(_p(g), _p(L['wt']), None, None, _p(dx), B, L['cout'], L['cout_pad'],
                                             L['cin'], _pad16(L['cin']), L['T'], 1, 0, s))
            prev = self.layers[i - 1]
            din = torch.empty(B, L['cin'], L['T'] // L['up'], device=self.device)
            _lib.check(lib.kym_dec_up_relu_adjoint(_p(dx), _p(acts[i]), _p(din), B, L['cin'], L['T'] // L['up'],
                                                    L['up'], int(prev['relu']), s))
            g = din
        return grads

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        lib = _lib.lib()
        s = _lib.stream_of(self.device)
        if self._adam is None:
            self._adam = [(torch.zeros_like(p), torch.zeros_like(p)) for p in self.parameters()]
        self.step_count += 1
        flat = [t for gw in grads for t in gw]
        for p, gr, (m, v) in zip(self.parameters(), flat, self._adam):
            _lib.check(lib.kym_dec_adam(_p(p), _p(gr), _p(m), _p(v), p.numel(), lr, b1, b2, eps, self.step_count, s))
        self._pack()
