"""Scattering1D as a torch.nn.Module on the MI355X HIP path.

Drop-in for kymatio.torch.Scattering1D (kymatio/scattering1d/frontend/
torch_frontend.py:8-83 and base_frontend.py:16-522): same constructor and
validation messages, same filter bank and float32 buffers ``tensor{n}``
(phi levels, psi1 levels, psi2 levels), same padding and border indices,
``meta()`` / ``output_size()``, and outputs 'array' (batch + (C, T)),
'list' (path dicts sorted by (order, n)) or 'dict' (keyed by n).  The
transform runs the fused pipeline of libkymatio_hip.so.
"""
import ctypes
import math
import numbers
from collections import Counter
from warnings import warn

import numpy as np
import torch

from ... import _lib
from ...frontend.base_frontend import ScatteringBase
from ...frontend.torch_frontend import ScatteringTorch
from ..filter_bank import anden_generator, compute_temporal_support, gauss_1d, scattering_filter_factory
from ..utils import compute_border_indices, compute_padding, parse_T

_MODE = {'local': 0, False: 1, 'global': 2}


class Scat1DPlanHandle:
    """Python handle of a kym_scat1d plan."""

    def __init__(self, module, max_order, dtype, lds_limit=0):
        lib = _lib.lib()
        m = module
        i32 = lambda v: (ctypes.c_int * max(len(v), 1))(*v)
        p1 = m.psi1_f
        p2 = m.psi2_f
        nind = max(m.ind_start.keys()) + 1
        h = ctypes.c_void_p()
        _lib.check(lib.kym_scat1d_create(
            m.shape[0], m._N_padded, m.pad_left, m.log2_stride, _MODE[m.average], m._oversampling, max_order,
            len(p1), i32([p['j'] for p in p1]), i32([len(p['levels']) for p in p1]),
            len(p2), i32([p['j'] for p in p2]), i32([len(p['levels']) for p in p2]),
            len(m.phi_f['levels']), i32([m.ind_start[i] for i in range(nind)]),
            i32([m.ind_end[i] for i in range(nind)]), nind, _lib.DTYPE_CODE[dtype], int(lds_limit),
            ctypes.byref(h)))
        self._h, self._lib = h, lib
        info = (ctypes.c_longlong * 7)()
        _lib.check(lib.kym_scat1d_info(h, info))
        (self.C, self.out_row, self.n_filters, self.filter_elems, self.ws_bytes_per_signal, self.n_launches,
         self.train_ws_bytes_per_signal) = list(info)
        fo = (ctypes.c_longlong * (self.n_filters + 1))()
        pl = (ctypes.c_longlong * self.C)()
        _lib.check(lib.kym_scat1d_layout(h, fo, pl))
        self.filter_offsets, self.path_len = list(fo), list(pl)

    def set_streams(self, streams):
        """Inference chunks in flight (0 = default 2; kym_scat1d_set_streams)."""
        _lib.check(self._lib.kym_scat1d_set_streams(self._h, int(streams)))

    def forward(self, x, filt, out, batch, ws):
        _lib.check(self._lib.kym_scat1d_forward(
            self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(filt.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), int(batch), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
            _lib.stream_of(x.device)))

    def forward_train(self, x, filt, out, batch, ws):
        _lib.check(self._lib.kym_scat1d_forward_train(
            self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(filt.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), int(batch), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
            _lib.stream_of(x.device)))

    def backward(self, dS, filt, gx, batch, ws):
        _lib.check(self._lib.kym_scat1d_backward(
            self._h, ctypes.c_void_p(dS.data_ptr()), ctypes.c_void_p(filt.data_ptr()),
            ctypes.c_void_p(gx.data_ptr()), int(batch), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
            _lib.stream_of(dS.device)))

    def __del__(self):
        try:
            if self._h:
                self._lib.kym_scat1d_destroy(self._h)
                self._h = None
        except Exception:
            pass


class ScatteringTorch1D(ScatteringTorch, ScatteringBase):
    def __init__(self, J, shape, Q=1, T=None, stride=None, max_order=2, oversampling=0, out_type='array',
                 backend='torch'):
        ScatteringTorch.__init__(self)
        self.J, self.shape, self._Q, self._T, self._stride = J, shape, Q, T, stride
        self.max_order, self.out_type, self.backend = max_order, out_type, backend
        self._oversampling = oversampling
        self._reduction = np.sum
        self._instantiate_backend('kymatio_amd.scattering1d.backend.')
        self.build()
        self.create_filters()
        self.register_filters()
        self._plans = {}
        self._packed = {}
        self.signals_per_chunk = None
        self.lds_limit = 0
        self.streams = 0  # inference chunks in flight on the device (0 = the library's default, 2)

    # ------------------------------------------------------------ setup
    def build(self):
        """kymatio/scattering1d/frontend/base_frontend.py:37-98."""
        self.r_psi = math.sqrt(0.5)
        self.sigma0 = 0.1
        self.alpha = 5.
        if isinstance(self.shape, numbers.Integral):
            self.shape = (self.shape,)
        elif isinstance(self.shape, tuple):
            if len(self.shape) > 1:
                raise ValueError('If shape is specified as a tuple, it must have exactly one element. '
                                 'Got: {}'.format(self.shape))
        else:
            raise ValueError('shape must be an integer or a 1-tuple. Got: {}'.format(self.shape))
        N_input = self.shape[0]
        if self._oversampling is None:
            pass
        elif self._oversampling < 0:
            raise ValueError('oversampling must be nonnegative. Got: {}'.format(self._oversampling))
        elif not isinstance(self._oversampling, numbers.Integral):
            raise ValueError('oversampling must be integer. Got: {}'.format(self._oversampling))
        elif self._oversampling > 0:
            warn('oversampling is deprecated in and will be removed in v0.5.Pass stride = 2**(J-oversampling) or '
                 'stride = 2**(log2(T)-oversampling) to retain current behavior.', DeprecationWarning)
            if self._stride not in [None, 2 ** (self.log2_T - self._oversampling - 1)]:
                raise ValueError('oversampling={} and stride={}are incompatible'.format(self._oversampling,
                                                                                      self._stride))
            self._stride = 2 ** (self.log2_T - self._oversampling)
        phi_f = gauss_1d(N_input, self.sigma0 / self.T)
        min_to_pad = 3 * compute_temporal_support(phi_f.reshape(1, -1), criterion_amplitude=1e-3)
        J_max_support = int(np.floor(np.log2(3 * N_input - 2)))
        J_pad = min(int(np.ceil(np.log2(N_input + 2 * min_to_pad))), J_max_support)
        self._N_padded = 2 ** J_pad
        self.pad_left, self.pad_right = compute_padding(self._N_padded, N_input)
        self.ind_start, self.ind_end = compute_border_indices(self.log2_T, self.J, self.pad_left,
                                                              self.pad_left + N_input)

    def create_filters(self):
        self.phi_f, self.psi1_f, self.psi2_f = scattering_filter_factory(
            self._N_padded, self.J, self.Q, self.T, self.filterbank, self._reduction)
        # no aliasing at first order, and j1 < j2 => xi1 > xi2 (frontend/base_frontend.py:17-30)
        assert all(p['xi'] < 0.5 / (2 ** p['j']) for p in self.psi1_f)
        assert all(p1['xi'] > p2['xi'] for p1 in self.psi1_f for p2 in self.psi2_f if p1['j'] < p2['j'])

    def register_filters(self):
        n = 0
        for bank in ([self.phi_f], self.psi1_f, self.psi2_f):
            for f in bank:
                for lv in range(len(f['levels'])):
                    f['levels'][lv] = torch.from_numpy(f['levels'][lv]).float().view(-1, 1)
                    self.register_buffer('tensor' + str(n), f['levels'][lv])
                    n += 1
        self._n_filters = n
        return n

    def load_filters(self):
        bufs = dict(self.named_buffers())
        n = 0
        for bank in ([self.phi_f], self.psi1_f, self.psi2_f):
            for f in bank:
                for lv in range(len(f['levels'])):
                    f['levels'][lv] = bufs['tensor' + str(n)]
                    n += 1
        return n

    # ------------------------------------------------------------ properties (base_frontend.py:232-299)
    @property
    def oversampling(self):
        warn('The attribute oversampling is deprecated and will be removed in v0.5. Pass stride = '
             '2**(J-oversampling) or stride = 2**(log2(T)-oversampling) to retain current behavior.',
             DeprecationWarning)
        return self._oversampling

    @property
    def average(self):
        return parse_T(self._T, self.J, self.shape[0], T_alias='T')[1]

    @property
    def filterbank(self):
        return (anden_generator, {'alpha': self.alpha, 'r_psi': self.r_psi, 'sigma0': self.sigma0})

    @property
    def log2_stride(self):
        if self._stride is None:
            return self.log2_T
        if not isinstance(self._stride, numbers.Integral):
            raise ValueError('stride must be integer. Got: {}'.format(self._stride))
        l2 = math.log2(self._stride)
        if math.floor(l2) != math.ceil(l2):
            raise ValueError('stride must be a power of two. Got: {}'.format(self._stride))
        if self.average in [False, 'global']:
            raise ValueError('stride={} is incompatible with T={}.'.format(self._stride, self._T))
        return int(l2)

    @property
    def log2_T(self):
        return int(math.floor(math.log2(self.T)))

    @property
    def Q(self):
        if np.any(np.array(self._Q) < 1):
            raise ValueError('Q must always be >= 1, got {}'.format(self._Q))
        if isinstance(self._Q, int):
            return (self._Q, 1)
        if isinstance(self._Q, tuple):
            if len(self._Q) == 1:
                return self._Q + (1,)
            if len(self._Q) == 2:
                return self._Q
            raise NotImplementedError('Q must be an integer, 1-tuple or 2-tuple. Scattering transforms beyond '
                                      'order 2 are not implemented. Got: {}'.format(self._Q))
        raise ValueError('Q must be an integer or a tuple')

    @property
    def stride(self):
        return 2 ** self.log2_stride

    @property
    def T(self):
        return parse_T(self._T, self.J, self.shape[0], T_alias='T')[0]

    # ------------------------------------------------------------ paths
    def _paths(self, max_order=None):
        """(order, n, j) of every output path in output order (sorted by (order, n))."""
        mo = self.max_order if max_order is None else max_order
        paths = [(0, (), ())]
        for n1, p in enumerate(self.psi1_f):
            paths.append((1, (n1,), (p['j'],)))
        if mo >= 2:
            for n1, p1 in enumerate(self.psi1_f):
                for n2, p2 in enumerate(self.psi2_f):
                    if p2['j'] > p1['j']:
                        paths.append((2, (n1, n2), (p1['j'], p2['j'])))
        return paths

    def meta(self):
        """kymatio/scattering1d/frontend/base_frontend.py:157-202."""
        paths = self._paths()
        meta = {'order': np.array([p[0] for p in paths]), 'key': [p[1] for p in paths]}
        meta['n'] = np.stack([np.append(p[1], (np.nan,) * (self.max_order - len(p[1]))) for p in paths])
        banks = (self.psi1_f, self.psi2_f)[:self.max_order]
        for key in ('xi', 'sigma', 'j'):
            meta[key] = meta['n'] * np.nan
            for order, bank in enumerate(banks):
                for n, psi in enumerate(bank):
                    meta[key][meta['n'][:, order] == n, order] = psi[key]
        return meta

    def output_size(self, detail=False):
        if detail:
            return tuple(Counter(self.meta()['order']).values())
        return len(self.meta()['key'])

    # ------------------------------------------------------------ transform
    def _check_runtime_args(self):
        if self.out_type not in ('array', 'dict', 'list'):
            raise ValueError("out_type must be one of 'array', 'dict', or 'list'. Got: {}".format(self.out_type))
        if not self.average and self.out_type == 'array':
            raise ValueError("Cannot convert to out_type='array' with T=0. Please set out_type to 'dict' or 'list'.")

    def _plan(self, device, dtype):
        key = (device.index, dtype, self.max_order)
        p = self._plans.get(key)
        if p is None:
            with torch.cuda.device(device):
                p = Scat1DPlanHandle(self, self.max_order, dtype, self.lds_limit)
                if self.streams:
                    p.set_streams(self.streams)
            self._plans[key] = p
        return p

    def _filter_buffers(self):
        bufs = dict(self.named_buffers())
        return [bufs['tensor' + str(i)] for i in range(self._n_filters)]

    def _packed_filters(self, plan, device, dtype):
        bufs = self._filter_buffers()
        key = (device.index, dtype)
        stamp = tuple((b.data_ptr(), b._version) for b in bufs)
        hit = self._packed.get(key)
        if hit is not None and hit[0] == stamp:
            return hit[1]
        offs = plan.filter_offsets
        packed = torch.empty(int(offs[-1]), dtype=dtype, device=device)
        for i, b in enumerate(bufs):
            packed[int(offs[i]):int(offs[i + 1])].copy_(b.reshape(-1))
        self._packed[key] = (stamp, packed)
        return packed

    def _forward_hip(self, x):
        device, dtype = x.device, x.dtype
        plan = self._plan(device, dtype)
        filt = self._packed_filters(plan, device, dtype)
        B = x.shape[0]
        out = torch.empty((B, plan.out_row), dtype=dtype, device=device)
        if B == 0:
            return out
        chunk = B if not self.signals_per_chunk else max(1, min(B, int(self.signals_per_chunk)))
        chunk = min(chunk, max(1, (512 << 20) // max(plan.ws_bytes_per_signal, 1)))
        ws = torch.empty(chunk * plan.ws_bytes_per_signal, dtype=torch.uint8, device=device)
        with torch.cuda.device(device):
            plan.forward(x, filt, out, B, ws)
        return out

    def scattering(self, x):
        self._check_runtime_args()
        if len(x.shape) < 1:
            raise ValueError('Input tensor x should have at least one axis, got {}'.format(len(x.shape)))
        bufs = self._filter_buffers()
        if x.dtype is not bufs[0].dtype:
            raise TypeError('Input and filter must be of the same dtype.')
        _lib.require_device(x)
        if not bufs[0].is_cuda:
            raise TypeError('Input must be on CPU.')
        if x.device != bufs[0].device:
            raise TypeError('Input and filter must be on the same GPU.')
        if x.shape[-1] != self.shape[0]:
            raise RuntimeError('Tensor must be of length %i.' % self.shape[0])
        batch_shape = x.shape[:-1]
        xr = x.reshape((-1, x.shape[-1])).contiguous()
        out = ScatteringFunction1D.apply(xr, self)
        plan = self._plan(x.device, x.dtype)
        paths = self._paths()
        if self.out_type == 'array':
            C = len(paths)
            return out.reshape(batch_shape + (C, plan.path_len[0]))
        offs = np.concatenate([[0], np.cumsum(plan.path_len)])
        coefs = [out[:, int(offs[c]):int(offs[c + 1])].reshape(batch_shape + (int(plan.path_len[c]),))
                 for c in range(len(paths))]
        if self.out_type == 'dict':
            return {p[1]: c for p, c in zip(paths, coefs)}
        return [{'coef': c, 'j': p[2], 'n': p[1], 'order': p[0]} for p, c in zip(paths, coefs)]


class ScatteringFunction1D(torch.autograd.Function):
    """Fused HIP forward; with a gradient needed, the training forward keeps the
    pre-modulus fields for the hand-written backward (kym_scat1d_backward)."""

    @staticmethod
    def forward(ctx, x, module):
        if not ctx.needs_input_grad[0]:
            return module._forward_hip(x)
        device, dtype = x.device, x.dtype
        plan = module._plan(device, dtype)
        filt = module._packed_filters(plan, device, dtype)
        B = x.shape[0]
        out = torch.empty((B, plan.out_row), dtype=dtype, device=device)
        ws = torch.empty(max(B, 1) * plan.train_ws_bytes_per_signal, dtype=torch.uint8, device=device)
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


__all__ = ['ScatteringTorch1D']
