"""Scattering2D as a torch.nn.Module on the MI355X HIP path.

Drop-in for kymatio.torch.Scattering2D
(kymatio/scattering2d/frontend/torch_frontend.py:8-111): same constructor,
buffers ``tensor{n}`` in the same order (phi levels, then each psi's levels),
same checks and messages, same output layout ``batch_shape + (C, h, w)`` and
path order.  ``scattering()`` runs the fused pipeline of libkymatio_hip.so
(one plan per device and dtype) instead of the per-op core loop; a
user-supplied backend object runs the reference core loop over its ops.
"""
import numpy as np
import torch

from ... import _lib
from ...frontend.base_frontend import ScatteringBase
from ...frontend.torch_frontend import ScatteringTorch
from ..filter_bank import filter_bank, leaf_taps
from ..utils import compute_padding, pad_amounts
from .plan import Scat2DPlanHandle


class ScatteringTorch2D(ScatteringTorch, ScatteringBase):
    def __init__(self, J, shape, L=8, max_order=2, pre_pad=False, backend='torch', out_type='array'):
        ScatteringTorch.__init__(self)
        self.J, self.shape, self.L, self.max_order = J, tuple(shape), L, max_order
        self.pre_pad, self.backend, self.out_type = pre_pad, backend, out_type
        self._instantiate_backend('kymatio_amd.scattering2d.backend.')
        M, N = self.shape
        if 2 ** J > M or 2 ** J > N:
            raise RuntimeError('The smallest dimension should be larger than 2^J.')
        self._M_padded, self._N_padded = compute_padding(M, N, J)
        self._pad_size = pad_amounts(M, N, J)
        filters = filter_bank(self._M_padded, self._N_padded, J, L)
        self.phi, self.psi = filters['phi'], filters['psi']
        self.register_filters()
        self._plans = {}
        self._packed = {}
        # images per launch chunk (workspace = chunk * per-image bytes); None = heuristic
        self.images_per_chunk = None
        # cap on LDS bytes per workgroup for the plans (0 = device maximum)
        self.lds_limit = 0
        # plan tuning choices (kym_scat2d_create_tuned; None = heuristic), see scattering2d/autotune.py
        self.tune = None

    # ------------------------------------------------------------ buffers
    def register_filters(self):
        n = 0
        for lev in self.phi['levels']:
            self.register_buffer('tensor' + str(n), torch.from_numpy(lev).unsqueeze(-1))
            n += 1
        for p in self.psi:
            for lev in p['levels']:
                self.register_buffer('tensor' + str(n), torch.from_numpy(lev).unsqueeze(-1))
                n += 1
        self._n_filters = n

    def _filter_buffers(self):
        bufs = dict(self.named_buffers())
        return [bufs['tensor' + str(i)] for i in range(self._n_filters)]

    def load_filters(self):
        """(phi, psi) dicts pointing at the module buffers
        (kymatio/scattering2d/frontend/torch_frontend.py:48-70)."""
        bufs = self._filter_buffers()
        n = 0
        phi = {k: v for k, v in self.phi.items() if k != 'levels'}
        phi['levels'] = []
        for _ in self.phi['levels']:
            phi['levels'].append(bufs[n])
            n += 1
        psis = []
        for p in self.psi:
            q = {k: v for k, v in p.items() if k != 'levels'}
            q['levels'] = []
            for _ in p['levels']:
                q['levels'].append(bufs[n])
                n += 1
            psis.append(q)
        return phi, psis

    # ------------------------------------------------------------ HIP path
    def _plan(self, device, dtype):
        key = (device.index, dtype, int(self.lds_limit), tuple(self.tune or ()))
        p = self._plans.get(key)
        if p is None:
            with torch.cuda.device(device):
                p = Scat2DPlanHandle(self.shape[0], self.shape[1], self.J, self.L, self.max_order,
                                     int(bool(self.pre_pad)), dtype, self.lds_limit, self.tune)
            self._plans[key] = p
        return p

    def _packed_filters(self, plan, device, dtype):
        bufs = self._filter_buffers()
        key = (device.index, dtype)
        stamp = tuple((b.data_ptr(), b._version) for b in bufs)
        hit = self._packed.get(key)
        if hit is not None and hit[0] == stamp:
            return hit[1]
        offs = plan.filter_offsets
        packed = torch.empty(int(plan.filter_elems), dtype=dtype, device=device)
        for i, b in enumerate(bufs):
            packed[int(offs[i]):int(offs[i + 1])].copy_(b.reshape(-1))
        desc = plan.leaf_taps_desc
        if desc[1] > 0:
            phi = [b.detach().cpu().numpy()[..., 0] for b in bufs[:self.J]]
            taps = leaf_taps(phi, desc)
            packed[int(desc[0]):int(desc[0] + desc[1])].copy_(torch.from_numpy(taps).to(dtype))
        self._packed[key] = (stamp, packed)
        return packed

    def _chunk(self, plan, batch):
        """Images per launch chunk: images_per_chunk if set, else as many as a
        256 MB workspace holds, balanced so that all chunks are equally full."""
        if self.images_per_chunk:
            return max(1, min(batch, int(self.images_per_chunk)))
        per = plan.ws_bytes_per_image
        cap = max(1, min(batch, (256 << 20) // max(per, 1)))
        nchunks = -(-batch // cap)
        return -(-batch // nchunks)

    def scattering(self, input):
        if not torch.is_tensor(input):
            raise TypeError('The input should be a PyTorch Tensor.')
        if len(input.shape) < 2:
            raise RuntimeError('Input tensor must have at least two dimensions.')
        if not input.is_contiguous():
            raise RuntimeError('Tensor must be contiguous.')
        if (input.shape[-1] != self.shape[-1] or input.shape[-2] != self.shape[-2]) and not self.pre_pad:
            raise RuntimeError('Tensor must be of spatial size (%i,%i).' % (self.shape[0], self.shape[1]))
        if (input.shape[-1] != self._N_padded or input.shape[-2] != self._M_padded) and self.pre_pad:
            raise RuntimeError('Padded tensor must be of spatial size (%i,%i).'
                               % (self._M_padded, self._N_padded))
        if self.out_type not in ('array', 'list'):
            raise RuntimeError("The out_type must be one of 'array' or 'list'.")
        if not getattr(self.backend, 'fused', False):
            return self._scattering_per_op(input)
        bufs = self._filter_buffers()
        if input.dtype is not bufs[0].dtype:
            raise TypeError('Input and filter must be of the same dtype.')
        _lib.require_device(input)
        if not bufs[0].is_cuda:
            raise TypeError('Input must be on CPU.')
        if input.device != bufs[0].device:
            raise TypeError('Input and filter must be on the same GPU.')
        batch_shape, sig = input.shape[:-2], input.shape[-2:]
        x = input.reshape((-1,) + tuple(sig))
        S = ScatteringFunction2D.apply(x, self)
        if self.out_type == 'array':
            return S.reshape(batch_shape + S.shape[-3:])
        out = []
        for ch, meta in enumerate(self._path_meta()):
            d = dict(meta)
            d['coef'] = S[:, ch].reshape(batch_shape + S.shape[-2:])
            out.append(d)
        return out

    def _forward_hip(self, x):
        device, dtype = x.device, x.dtype
        plan = self._plan(device, dtype)
        filt = self._packed_filters(plan, device, dtype)
        B = x.shape[0]
        out = torch.empty((B, plan.C, plan.oh, plan.ow), dtype=dtype, device=device)
        if B == 0:
            return out
        chunk = self._chunk(plan, B)
        ws = torch.empty(chunk * plan.ws_bytes_per_image, dtype=torch.uint8, device=device)
        with torch.cuda.device(device):
            plan.forward(x, filt, out, B, ws)
        return out

    def _path_meta(self):
        """(j, n, theta) of each output channel, in the reference's path order
        (kymatio/scattering2d/core/scattering2d.py:25-78)."""
        metas = [{'j': (), 'n': (), 'theta': ()}]
        for n1, p in enumerate(self.psi):
            metas.append({'j': (p['j'],), 'n': (n1,), 'theta': (p['theta'],)})
        if self.max_order >= 2:
            for n1, p1 in enumerate(self.psi):
                for n2, p2 in enumerate(self.psi):
                    if p2['j'] > p1['j']:
                        metas.append({'j': (p1['j'], p2['j']), 'n': (n1, n2),
                                      'theta': (p1['theta'], p2['theta'])})
        return metas

    def _scattering_per_op(self, input):
        from ..core.scattering2d import scattering2d
        phi, psi = self.load_filters()
        batch_shape, sig = input.shape[:-2], input.shape[-2:]
        x = input.reshape((-1,) + tuple(sig))
        if self.pre_pad:
            pad = lambda z: z.reshape(z.shape + (1,))
        else:
            pad = self.backend.Pad(self._pad_size, list(self.shape))
        S = scattering2d(x, pad, self.backend.unpad, self.backend, self.J, self.L, phi, psi,
                         self.max_order, self.out_type)
        if self.out_type == 'array':
            return S.reshape(batch_shape + S.shape[-3:])
        for s in S:
            s['coef'] = s['coef'].reshape(batch_shape + s['coef'].shape[-2:])
        return S


class ScatteringFunction2D(torch.autograd.Function):
    """Fused HIP forward; when a gradient is needed, the training forward keeps
    every node's pre-modulus field in a workspace that the hand-written
    backward (kym_scat2d_backward) consumes."""

    @staticmethod
    def forward(ctx, x, module):
        x = x.contiguous()
        if not ctx.needs_input_grad[0]:
            return module._forward_hip(x)
        device, dtype = x.device, x.dtype
        plan = module._plan(device, dtype)
        if plan.train_ws_bytes_per_image < 0:
            raise NotImplementedError('this configuration has no backward on the HIP path')
        filt = module._packed_filters(plan, device, dtype)
        B = x.shape[0]
        out = torch.empty((B, plan.C, plan.oh, plan.ow), dtype=dtype, device=device)
        ws = torch.empty(max(B, 1) * plan.train_ws_bytes_per_image, dtype=torch.uint8, device=device)
        if B:
            with torch.cuda.device(device):
                plan.forward_train(x, filt, out, B, ws)
        ctx.save_for_backward(filt, ws)
        ctx.plan, ctx.B, ctx.xshape = plan, B, x.shape
        return out

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad):
        filt, ws = ctx.saved_tensors
        g = grad.contiguous()
        gx = torch.empty(ctx.xshape, dtype=g.dtype, device=g.device)
        if ctx.B:
            with torch.cuda.device(g.device):
                ctx.plan.backward(g, filt, gx, ctx.B, ws)
        return gx, None


__all__ = ['ScatteringTorch2D']
