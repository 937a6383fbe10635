"""Oracle (test infrastructure only): Scattering2D forward, restated on the CPU.

Two restatements of the same path:

* ``scattering2d_torch`` -- op for op the reference *torch* backend on CPU
  tensors (FP32 "as reference", or FP64), differentiable through torch
  autograd exactly like the reference (the modulus uses the reference's
  stable sub-gradient).  Ops follow:
    pad        kymatio/scattering2d/backend/torch_backend.py:19-86 (ReflectionPad2d
               + the pad == size special case at :48-54, :80-83)
    rfft       :133-141 (zeros + copy, then fft2)
    cdgmm      kymatio/backend/torch_backend.py:147-219 (real filter => A * B)
    subsample  kymatio/scattering2d/backend/torch_backend.py:92-129 (view, mean(3), mean(1))
    ifft/irfft :143-155 (ifft2, real part)
    modulus    kymatio/backend/torch_backend.py:5-96, :137-141
    unpad      kymatio/scattering2d/backend/torch_backend.py:157-176
    core loop  kymatio/scattering2d/core/scattering2d.py:1-88
* ``scattering2d_numpy`` -- the NumPy backend twin
  (kymatio/scattering2d/backend/numpy_backend.py), used in FP64 as the
  high-precision reference the error of every order is stated against.
"""
import numpy as np
import torch

from . import filters2d


class _ModulusStable(torch.autograd.Function):
    """kymatio/backend/torch_backend.py:5-96: |z| with gradient 0 at z == 0."""

    @staticmethod
    def forward(ctx, x):
        out = (x[..., 0] * x[..., 0] + x[..., 1] * x[..., 1]).sqrt()
        ctx.save_for_backward(x, out)
        return out

    @staticmethod
    def backward(ctx, g):
        x, out = ctx.saved_tensors
        gi = x.mul(g.unsqueeze(-1)).div(out.unsqueeze(-1))
        gi.masked_fill_((out == 0).unsqueeze(-1), 0)
        return gi


def pad_torch(x, pad_size, input_size):
    """kymatio/scattering2d/backend/torch_backend.py:19-86."""
    p = list(pad_size)
    if p[0] == input_size[0]:
        p[0] -= 1
        p[1] -= 1
    if p[2] == input_size[1]:
        p[2] -= 1
        p[3] -= 1
    mod = torch.nn.ReflectionPad2d([p[2], p[3], p[0], p[1]])
    batch_shape, sig = x.shape[:-2], x.shape[-2:]
    x = mod(x.reshape((-1, 1) + sig))
    if pad_size[0] == input_size[0]:
        x = torch.cat([x[:, :, 1, :].unsqueeze(2), x,
                       x[:, :, x.shape[2] - 2, :].unsqueeze(2)], 2)
    if pad_size[2] == input_size[1]:
        x = torch.cat([x[:, :, :, 1].unsqueeze(3), x,
                       x[:, :, :, x.shape[3] - 2].unsqueeze(3)], 3)
    return x.reshape(batch_shape + x.shape[-2:] + (1,))


def _fft(x):
    return torch.view_as_real(torch.fft.fft2(torch.view_as_complex(x)))


def _ifft(x):
    return torch.view_as_real(torch.fft.ifft2(torch.view_as_complex(x)))


def rfft_torch(x):
    xr = torch.zeros(x.shape[:-1] + (2,), dtype=x.dtype, device=x.device)
    xr[..., 0] = x[..., 0]
    return _fft(xr)


def irfft_torch(x):
    return torch.fft.ifft2(torch.view_as_complex(x)).real[..., None]


def ifft_torch(x):
    return _ifft(x)


def cdgmm_torch(a, b):
    return a * b


def subsample_torch(x, k):
    bs, sig = x.shape[:-3], x.shape[-3:]
    x = x.view((-1,) + sig)
    y = x.view(-1, k, x.shape[1] // k, k, x.shape[2] // k, 2)
    out = y.mean(3, keepdim=False).mean(1, keepdim=False)
    return out.reshape(bs + out.shape[-3:])


def modulus_torch(x):
    return _ModulusStable.apply(x)[..., None]


def unpad_torch(x):
    x = x[..., 1:-1, 1:-1, :]
    return x.reshape(x.shape[:-1])


def padding_amounts(M, N, J):
    """kymatio/scattering2d/frontend/base_frontend.py:19-32."""
    Mp, Np = filters2d.compute_padding(M, N, J)
    return (Mp, Np), [(Mp - M) // 2, (Mp - M + 1) // 2, (Np - N) // 2, (Np - N + 1) // 2]


def core_torch(x, pad, J, phi, psi, max_order=2, out_type='array'):
    """kymatio/scattering2d/core/scattering2d.py:1-88."""
    out0, out1, out2 = [], [], []
    U_r = pad(x)
    U_0_c = rfft_torch(U_r)
    U_1_c = subsample_torch(cdgmm_torch(U_0_c, phi['levels'][0]), 2 ** J)
    S_0 = unpad_torch(irfft_torch(U_1_c))
    out0.append({'coef': S_0, 'j': (), 'n': (), 'theta': ()})
    for n1 in range(len(psi)):
        j1, t1 = psi[n1]['j'], psi[n1]['theta']
        U_1_c = cdgmm_torch(U_0_c, psi[n1]['levels'][0])
        if j1 > 0:
            U_1_c = subsample_torch(U_1_c, 2 ** j1)
        U_1_c = rfft_torch(modulus_torch(ifft_torch(U_1_c)))
        S_1_c = subsample_torch(cdgmm_torch(U_1_c, phi['levels'][j1]), 2 ** (J - j1))
        out1.append({'coef': unpad_torch(irfft_torch(S_1_c)), 'j': (j1,), 'n': (n1,),
                     'theta': (t1,)})
        if max_order < 2:
            continue
        for n2 in range(len(psi)):
            j2, t2 = psi[n2]['j'], psi[n2]['theta']
            if j2 <= j1:
                continue
            U_2_c = subsample_torch(cdgmm_torch(U_1_c, psi[n2]['levels'][j1]), 2 ** (j2 - j1))
            U_2_c = rfft_torch(modulus_torch(ifft_torch(U_2_c)))
            S_2_c = subsample_torch(cdgmm_torch(U_2_c, phi['levels'][j2]), 2 ** (J - j2))
            out2.append({'coef': unpad_torch(irfft_torch(S_2_c)), 'j': (j1, j2),
                         'n': (n1, n2), 'theta': (t1, t2)})
    out = out0 + out1 + out2
    if out_type == 'array':
        return torch.stack([o['coef'] for o in out], -3)
    return out


class Scattering2DTorchCPU:
    """Mirror of ScatteringTorch2D (kymatio/scattering2d/frontend/torch_frontend.py:8-111)
    on the CPU, with filters built by the oracle's filter bank."""

    def __init__(self, J, shape, L=8, max_order=2, pre_pad=False, dtype=torch.float32,
                 filters=None):
        M, N = shape
        if 2 ** J > M or 2 ** J > N:
            raise RuntimeError('The smallest dimension should be larger than 2^J.')
        self.J, self.L, self.shape, self.max_order, self.pre_pad = J, L, shape, max_order, pre_pad
        (self.Mp, self.Np), self.pads = padding_amounts(M, N, J)
        fb = filters if filters is not None else filters2d.filter_bank(self.Mp, self.Np, J, L)
        self.filters_np = fb
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).unsqueeze(-1).to(dtype)
        self.phi = {'levels': [t(v) for v in fb['phi']['levels']], 'j': fb['phi']['j']}
        self.psi = [{'levels': [t(v) for v in p['levels']], 'j': p['j'], 'theta': p['theta']}
                    for p in fb['psi']]

    def __call__(self, x, out_type='array'):
        batch_shape, sig = x.shape[:-2], x.shape[-2:]
        x = x.reshape((-1,) + sig)
        if self.pre_pad:
            pad = lambda z: z.reshape(z.shape + (1,))
        else:
            pad = lambda z: pad_torch(z, self.pads, list(self.shape))
        S = core_torch(x, pad, self.J, self.phi, self.psi, self.max_order, out_type)
        if out_type == 'array':
            return S.reshape(batch_shape + S.shape[-3:])
        for s in S:
            s['coef'] = s['coef'].reshape(batch_shape + s['coef'].shape[-2:])
        return S


def scattering2d_numpy(x, J, L=8, max_order=2, filters=None, dtype=np.float64):
    """NumPy backend twin (kymatio/scattering2d/backend/numpy_backend.py,
    kymatio/backend/numpy_backend.py) of the same core loop, in ``dtype``."""
    import scipy.fft as sfft
    x = np.asarray(x, dtype)
    batch_shape, (M, N) = x.shape[:-2], x.shape[-2:]
    (Mp, Np), pads = padding_amounts(M, N, J)
    fb = filters if filters is not None else filters2d.filter_bank(Mp, Np, J, L)
    x = x.reshape((-1, M, N))
    U_r = np.pad(x, ((0, 0), (pads[0], pads[1]), (pads[2], pads[3])), mode='reflect')

    def sub(z, k):
        y = z.reshape(-1, k, z.shape[1] // k, k, z.shape[2] // k)
        return y.mean(axis=(1, 3))

    def unpad(z):
        return z[..., 1:-1, 1:-1]

    out = []
    U0 = sfft.fft2(U_r)
    phi, psi = fb['phi'], fb['psi']
    out.append(unpad(sfft.ifft2(sub(U0 * phi['levels'][0], 2 ** J)).real))
    o2 = []
    for n1 in range(len(psi)):
        j1 = psi[n1]['j']
        U1 = U0 * psi[n1]['levels'][0]
        if j1 > 0:
            U1 = sub(U1, 2 ** j1)
        U1 = sfft.fft2(np.abs(sfft.ifft2(U1)))
        out.append(unpad(sfft.ifft2(sub(U1 * phi['levels'][j1], 2 ** (J - j1))).real))
        if max_order < 2:
            continue
        for n2 in range(len(psi)):
            j2 = psi[n2]['j']
            if j2 <= j1:
                continue
            U2 = sub(U1 * psi[n2]['levels'][j1], 2 ** (j2 - j1))
            U2 = sfft.fft2(np.abs(sfft.ifft2(U2)))
            o2.append(unpad(sfft.ifft2(sub(U2 * phi['levels'][j2], 2 ** (J - j2))).real))
    S = np.stack(out + o2, axis=-3)
    return S.reshape(batch_shape + S.shape[-3:])


def channel_orders(J, L, max_order=2):
    """Order of each output channel in the reference's path order (S0, S1..., S2...)."""
    orders = [0] + [1] * (J * L)
    if max_order >= 2:
        orders += [2] * (L * L * J * (J - 1) // 2)
    return np.array(orders)
