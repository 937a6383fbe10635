"""Oracle (test infrastructure only): Scattering1D forward, restated on the CPU.

``Scattering1DTorchCPU`` runs, op for op, the reference torch backend on CPU
tensors, differentiable through torch autograd:
  pad        kymatio/scattering1d/backend/torch_backend.py:50-82 (F.pad reflect)
  rfft/ifft/irfft/cfft  :108-141
  subsample  :18-48 (view (k, N/k, 2), mean over k)
  cdgmm      kymatio/backend/torch_backend.py:147-219 (real filter => A * B)
  modulus    kymatio/backend/torch_backend.py:5-96
  unpad      kymatio/scattering1d/backend/torch_backend.py:84-106
  average_global :143-148
  core       kymatio/scattering1d/core/scattering1d.py:2-107
  frontend   kymatio/scattering1d/frontend/base_frontend.py:106-155 (unpad at
             ind_start[res]:ind_end[res], sort by (order, n), stack on dim -2)
Filters are the float64 factory output cast to float32, as the torch frontend
does (kymatio/scattering1d/frontend/torch_frontend.py:30-56).
"""
import numpy as np
import torch
import torch.nn.functional as F

from . import filters1d
from .scat2d import _ModulusStable


def _fft(x):
    return torch.view_as_real(torch.fft.fft(torch.view_as_complex(x)))


def _ifft(x):
    return torch.view_as_real(torch.fft.ifft(torch.view_as_complex(x)))


def rfft(x):
    xr = torch.zeros(x.shape[:-1] + (2,), dtype=x.dtype, device=x.device)
    xr[..., 0] = x[..., 0]
    return _fft(xr)


def irfft(x):
    return torch.fft.ifft(torch.view_as_complex(x)).real[..., None]


def subsample(x, k):
    N = x.shape[-2]
    return x.view(x.shape[:-2] + (k, N // k, 2)).mean(dim=-3)


def modulus(x):
    return _ModulusStable.apply(x)[..., None]


def pad(x, left, right):
    if left >= x.shape[-1] or right >= x.shape[-1]:
        raise ValueError('Indefinite padding size (larger than tensor).')
    return F.pad(x, (left, right), mode='reflect')[..., None]


def unpad(x, i0, i1):
    return x.reshape(x.shape[:-1])[..., i0:i1]


def core(U_0, filters, log2_stride, average_local):
    """Generator; kymatio/scattering1d/core/scattering1d.py:2-107."""
    U_0_hat = rfft(U_0)
    phi = filters[0]
    if average_local:
        S_0 = irfft(subsample(U_0_hat * phi['levels'][0], 2 ** log2_stride))
        yield {'coef': S_0, 'j': (), 'n': ()}
    else:
        yield {'coef': U_0, 'j': (), 'n': ()}
    psi1 = filters[1]
    for n1 in range(len(psi1)):
        j1 = psi1[n1]['j']
        k1 = min(j1, log2_stride) if average_local else j1
        U_1_hat = subsample(U_0_hat * psi1[n1]['levels'][0], 2 ** k1)
        U_1_m = modulus(_ifft(U_1_hat))
        if average_local or len(filters) > 2:
            U_1_hat = rfft(U_1_m)
        if average_local:
            k1_J = max(log2_stride - k1, 0)
            S_1 = irfft(subsample(U_1_hat * phi['levels'][k1], 2 ** k1_J))
            yield {'coef': S_1, 'j': (j1,), 'n': (n1,)}
        else:
            yield {'coef': U_1_m, 'j': (j1,), 'n': (n1,)}
        if len(filters) > 2:
            psi2 = filters[2]
            for n2 in range(len(psi2)):
                j2 = psi2[n2]['j']
                if j2 > j1:
                    sub2 = min(j2, log2_stride) if average_local else j2
                    k2 = max(sub2 - k1, 0)
                    U_2_hat = subsample(U_1_hat * psi2[n2]['levels'][k1], 2 ** k2)
                    U_2_m = modulus(_ifft(U_2_hat))
                    if average_local:
                        U_2_hat = rfft(U_2_m)
                        k2_T = max(log2_stride - sub2, 0)
                        S_2 = irfft(subsample(U_2_hat * phi['levels'][k1 + k2], 2 ** k2_T))
                        yield {'coef': S_2, 'j': (j1, j2), 'n': (n1, n2)}
                    else:
                        yield {'coef': U_2_m, 'j': (j1, j2), 'n': (n1, n2)}


class Scattering1DTorchCPU:
    """Mirror of ScatteringTorch1D.scattering on the CPU."""

    def __init__(self, J, shape, Q=1, T=None, stride=None, max_order=2, oversampling=0,
                 out_type='array', dtype=torch.float32):
        self.J, self.N, self.max_order, self.out_type = J, shape, max_order, out_type
        self.cfg = filters1d.setup(J, shape, Q, T, stride, oversampling)
        c = self.cfg
        t = lambda a: torch.from_numpy(a).float().view(-1, 1).to(dtype)
        self.phi = dict(c['phi'], levels=[t(v) for v in c['phi']['levels']])
        self.psi1 = [dict(p, levels=[t(v) for v in p['levels']]) for p in c['psi1']]
        self.psi2 = [dict(p, levels=[t(v) for v in p['levels']]) for p in c['psi2']]

    def __call__(self, x):
        c = self.cfg
        batch_shape, sig = x.shape[:-1], x.shape[-1:]
        x = x.reshape((-1, 1) + sig)
        U_0 = pad(x, c['pad_left'], c['pad_right'])
        filters = [self.phi, self.psi1, self.psi2][:1 + self.max_order]
        S = [] if self.out_type in ('array', 'list') else {}
        for path in core(U_0, filters, c['log2_stride'], c['average'] == 'local'):
            path['order'] = len(path['n'])
            if c['average'] == 'local':
                res = c['log2_stride']
            elif path['order'] > 0:
                res = max(path['j'][-1] - c['oversampling'], 0)
            else:
                res = 0
            if c['average'] == 'global':
                path['coef'] = torch.sum(path['coef'], axis=-2, keepdims=True)
            else:
                path['coef'] = unpad(path['coef'], c['ind_start'][res], c['ind_end'][res])
            path['coef'] = path['coef'].reshape(batch_shape + path['coef'].shape[-1:])
            if isinstance(S, list):
                S.append(path)
            else:
                S[path['n']] = path['coef']
        if isinstance(S, dict):
            return S
        S.sort(key=lambda p: (p['order'], p['n']))
        if self.out_type == 'array':
            return torch.stack([p['coef'] for p in S], dim=-2)
        return S
