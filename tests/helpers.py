"""Shared test helpers: golden fixtures, packed filters, error metrics."""
import ctypes
import os

import numpy as np

from kymatio_amd.scattering2d.filter_bank import leaf_taps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, 'tests', 'golden')

# Parity bar for FP32 (stated in the workload definition): per output channel,
# max |ours - ref| <= 1e-5 * max |ref| over that channel (relative to the
# channel's scale; element-wise relative error is ill-posed for coefficients
# near zero).
RTOL_CHANNEL = 1e-5


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def channel_rel_err(ours, ref):
    """max over (batch, channel) of max|ours-ref| / max|ref| of that channel."""
    ours = np.asarray(ours, np.float64)
    ref = np.asarray(ref, np.float64)
    c_axis = ref.ndim - 3 if ref.ndim >= 3 else 0
    lead = int(np.prod(ref.shape[:c_axis + 1]))
    o = ours.reshape(lead, -1)
    r = ref.reshape(lead, -1)
    scale = np.abs(r).max(axis=1)
    scale[scale == 0] = 1.0
    return float((np.abs(o - r).max(axis=1) / scale).max())


def pack_filters_2d(fb, dtype=np.float32):
    arrs = list(fb['phi']['levels']) + [lev for p in fb['psi'] for lev in p['levels']]
    return np.concatenate([np.asarray(a).reshape(-1) for a in arrs]).astype(dtype)


def golden_2d_cases():
    G = golden('golden_scat2d.npz')
    tags = ['cfg1', 'cfg1o1', 'cfg4', 'odd31', 'odd33', 'rect37x32', 'sq32J5', 'rect40x24', 'prepad', 'L6', 'cfg2']
    out = []
    for tag in tags:
        kw = {}
        for k in ('pre_pad', 'L', 'max_order'):
            if tag + '_' + k in G.files:
                kw[k] = int(G[tag + '_' + k])
        out.append((tag, G[tag + '_x'], int(G[tag + '_J']), tuple(int(v) for v in G[tag + '_shape']), kw,
                    G[tag + '_Sx']))
    return out


class Emu:
    """ctypes binding of libkymatio_emu.so (host-only harness running the GPU
    workgroup bodies single-threaded)."""

    def __init__(self):
        path = os.path.join(ROOT, 'kymatio_amd', 'libkymatio_emu.so')
        self.lib = ctypes.CDLL(path)
        f = self.lib.kym_emu_scat2d_forward
        f.argtypes = [ctypes.c_int] * 7 + [ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_longlong, ctypes.c_longlong,
                                           ctypes.c_void_p]
        f.restype = ctypes.c_int
        self.lib.kym_emu_last_error.restype = ctypes.c_char_p

    def with_taps(self, filters_packed, J, shape, L=8, max_order=2, pre_pad=0, lds_limit=160 * 1024):
        """The packed filters followed by the plan's low-pass taps
        (filter_bank.leaf_taps from the phi levels at the buffer's start)."""
        dt = filters_packed.dtype
        code = 0 if dt == np.float32 else 1
        f = self.lib.kym_emu_scat2d_taps
        f.argtypes = [ctypes.c_int] * 7 + [ctypes.c_longlong, ctypes.c_void_p, ctypes.c_int]
        f.restype = ctypes.c_int
        n = 4 + 10 * max(J, 1)
        desc = (ctypes.c_longlong * n)()
        rc = f(shape[0], shape[1], J, L, max_order, pre_pad, code, lds_limit, desc, n)
        if rc != 0:
            raise RuntimeError(self.lib.kym_emu_last_error().decode())
        desc = list(desc)[:4 + 10 * J]
        if desc[1] == 0:
            return filters_packed
        Mp, Np = desc[2], desc[3]
        phi, o = [], 0
        for r in range(J):
            h, w = Mp >> r, Np >> r
            phi.append(filters_packed[o:o + h * w].reshape(h, w))
            o += h * w
        out = np.zeros(desc[0] + desc[1], dt)
        out[:desc[0]] = filters_packed[:desc[0]]
        out[desc[0]:] = leaf_taps(phi, desc).astype(dt)
        return out

    def scat2d(self, x, J, shape, filters_packed, L=8, max_order=2, pre_pad=0, lds_limit=160 * 1024,
               chunk=0, tune=None):
        dt = filters_packed.dtype
        code = 0 if dt == np.float32 else 1
        M, N = shape
        B = x.shape[0]
        info = (ctypes.c_longlong * 9)()
        rc = self.lib.kym_emu_scat2d_forward(M, N, J, L, max_order, pre_pad, code, lds_limit, None, None, None,
                                             B, 0, info)
        if rc != 0:
            raise RuntimeError(self.lib.kym_emu_last_error().decode())
        out = np.zeros((B, info[0], info[1], info[2]), dt)
        xx = np.ascontiguousarray(x, dt)
        filters_packed = self.with_taps(filters_packed, J, shape, L, max_order, pre_pad, lds_limit)
        if tune:
            f = self.lib.kym_emu_scat2d_forward_tuned
            f.argtypes = self.lib.kym_emu_scat2d_forward.argtypes + [ctypes.c_void_p, ctypes.c_int]
            f.restype = ctypes.c_int
            t = (ctypes.c_int * len(tune))(*tune)
            rc = f(M, N, J, L, max_order, pre_pad, code, lds_limit, xx.ctypes.data, filters_packed.ctypes.data,
                   out.ctypes.data, B, chunk, info, t, len(tune))
        else:
            rc = self.lib.kym_emu_scat2d_forward(M, N, J, L, max_order, pre_pad, code, lds_limit, xx.ctypes.data,
                                                 filters_packed.ctypes.data, out.ctypes.data, B, chunk, info)
        if rc != 0:
            raise RuntimeError(self.lib.kym_emu_last_error().decode())
        return out, {'launches': info[3], 'big_nodes': info[4], 'spec_launches': info[5], 'pair_launches': info[6],
                     'bwd_spec_launches': info[7], 'col_pair_launches': info[8]}


def emu_scat1d(emu, module, x, dtype=np.float32, dS=None, nbwd=1):
    """Run the 1D bodies in the CPU harness with the setup of a (CPU)
    Scattering1D module; with dS, also the backward (returns S, C, dL/dx)."""
    m = module
    modes = {'local': 0, False: 1, 'global': 2}
    i32 = lambda v: (ctypes.c_int * max(len(v), 1))(*v)
    f = emu.lib.kym_emu_scat1d_forward
    f.argtypes = ([ctypes.c_int] * 8 + [ctypes.c_void_p] * 2 + [ctypes.c_int] + [ctypes.c_void_p] * 2
                  + [ctypes.c_int] + [ctypes.c_void_p] * 2 + [ctypes.c_int, ctypes.c_int]
                  + [ctypes.c_void_p] * 3 + [ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_int])
    f.restype = ctypes.c_int
    p1, p2 = m.psi1_f, m.psi2_f
    nind = max(m.ind_start.keys()) + 1
    bufs = [lev.numpy().reshape(-1) for bank in ([m.phi_f], p1, p2) for flt in bank for lev in flt['levels']]
    packed = np.concatenate(bufs).astype(dtype)
    args = [m.shape[0], m._N_padded, m.pad_left, m.log2_stride, modes[m.average], m._oversampling, m.max_order,
            len(p1), i32([p['j'] for p in p1]), i32([len(p['levels']) for p in p1]),
            len(p2), i32([p['j'] for p in p2]), i32([len(p['levels']) for p in p2]),
            len(m.phi_f['levels']), i32([m.ind_start[i] for i in range(nind)]),
            i32([m.ind_end[i] for i in range(nind)]), nind, 0 if dtype == np.float32 else 1]
    info = (ctypes.c_longlong * 5)()
    B = x.shape[0]
    rc = f(*args, None, None, None, B, info, None, None, 0)
    if rc != 0:
        raise RuntimeError(emu.lib.kym_emu_last_error().decode())
    emu.last_spec_launches1d = info[4]  # forward launches on compile-time node kernels
    out = np.zeros((B, info[1]), dtype)
    xx = np.ascontiguousarray(x, dtype)
    if dS is None:
        rc = f(*args, xx.ctypes.data, packed.ctypes.data, out.ctypes.data, B, info, None, None, 0)
        if rc != 0:
            raise RuntimeError(emu.lib.kym_emu_last_error().decode())
        return out, info[0]
    dd = np.ascontiguousarray(dS, dtype).reshape(B, -1)
    gx = np.zeros_like(xx)
    rc = f(*args, xx.ctypes.data, packed.ctypes.data, out.ctypes.data, B, info, dd.ctypes.data, gx.ctypes.data,
           nbwd)
    if rc != 0:
        raise RuntimeError(emu.lib.kym_emu_last_error().decode())
    return out, info[0], gx


def emu_train2d(emu, x, J, shape, filters_packed, dS, L=8, max_order=2, pre_pad=0, lds_limit=160 * 1024, nbwd=1,
                tune=None):
    """Training forward + hand-written backward (nbwd times on the same saved
    state) of the 2D bodies on the CPU harness, with the plan tuning choices
    `tune` if given; returns (S, dL/dx of the last backward) for the upstream
    gradient dS."""
    f = emu.lib.kym_emu_scat2d_train_tuned
    f.argtypes = [ctypes.c_int] * 7 + [ctypes.c_longlong] + [ctypes.c_void_p] * 5 + [ctypes.c_longlong, ctypes.c_int,
                                                                                      ctypes.c_void_p, ctypes.c_int]
    f.restype = ctypes.c_int
    tv = (ctypes.c_int * 7)(*(list(tune or []) + [0] * 7)[:7])
    dt = filters_packed.dtype
    filters_packed = emu.with_taps(filters_packed, J, shape, L, max_order, pre_pad, lds_limit)
    xx = np.ascontiguousarray(x, dt)
    dd = np.ascontiguousarray(dS, dt)
    out = np.zeros(dS.shape, dt)
    gx = np.zeros(xx.shape, dt)
    rc = f(shape[0], shape[1], J, L, max_order, pre_pad, 0 if dt == np.float32 else 1, lds_limit, xx.ctypes.data,
           filters_packed.ctypes.data, out.ctypes.data, dd.ctypes.data, gx.ctypes.data, x.shape[0], nbwd, tv,
           len(tune or []))
    if rc != 0:
        raise RuntimeError(emu.lib.kym_emu_last_error().decode())
    return out, gx
