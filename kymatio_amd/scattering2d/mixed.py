"""Scattering2D over one batch of images of several sizes.

The reference's modules take one image size each
(kymatio/scattering2d/frontend/torch_frontend.py:9-21).  MixedScattering2D
keeps one such module per size (J chosen per size, by default
J = log2(side) - 3 so every size gives the same 8 x 8 output grid) and runs a
whole mixed batch as ONE replay of a HIP graph (kym_scat2d_graph_*): each
size's launch chain is an independent branch, so the sizes share the GPU
instead of running one after another.  Plans are autotuned per (size, count)
on first use (scattering2d/autotune.py).

The graph is laid out for a maximum image count per size
(kym_scat2d_graph_create_counted) and its kernels read the actual counts
from a small device array at every replay, so batches with other counts
replay the same graph: it is rebuilt only when a size is new or its count
exceeds the maximum (which then grows).  The graph is bound to fixed device
buffers of the maximum sizes; inputs are copied into them and outputs copied
out on every call.
"""
import ctypes
import math

import torch

from .. import _lib
from . import autotune as _autotune
from .frontend.torch_frontend import ScatteringTorch2D


def default_J(side):
    """J = log2(side) - 3 (the mixed-size workload's rule: 8 x 8 outputs)."""
    return int(round(math.log2(side))) - 3


class MixedScattering2D:
    def __init__(self, L=8, max_order=2, J_of_side=None, device='cuda', autotune=True, tune_cache=None,
                 ws_budget=256 << 20):
        self.L, self.max_order = L, max_order
        self.J_of_side = J_of_side or default_J
        self.device = torch.device(device)
        self.autotune = autotune
        self.tune_cache = tune_cache
        self.ws_budget = ws_budget
        self._mods = {}
        self._state = None  # (per-size maximum counts, graph handle, per-size buffers)
        self.tuning = {}    # side -> chosen plan parameters
        self.builds = 0     # graphs built so far (a replay with new counts within the maxima builds none)
        self._counts = None

    def module(self, side):
        """The ScatteringTorch2D used for square images of this side."""
        m = self._mods.get(side)
        if m is None:
            m = ScatteringTorch2D(self.J_of_side(side), (side, side), L=self.L, max_order=self.max_order)
            m = m.to(self.device)
            self._mods[side] = m
        return m

    def _release(self):
        if self._state is not None:
            _lib.lib().kym_scat2d_graph_destroy(self._state[1])
            self._state = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    @staticmethod
    def _headroom(n):
        """Maximum count the graph is laid out for when a size first needs n."""
        return max(n + 16, -(-5 * n // 4))

    def _build(self, maxima):
        self._release()
        lib = _lib.lib()
        sides = sorted(maxima)
        bufs = []
        for side in sides:
            n = maxima[side]
            m = self.module(side)
            if self.autotune:
                choice, _ = _autotune.tune(m, n, cache_path=self.tune_cache)
                self.tuning[side] = choice
            plan = m._plan(self.device, torch.float32)
            filt = m._packed_filters(plan, self.device, torch.float32)
            # the autotuned images per chunk, else as many as ws_budget holds
            if m.images_per_chunk:
                chunk = m._chunk(plan, n)
            else:
                chunk = max(1, min(n, self.ws_budget // max(plan.ws_bytes_per_image, 1)))
            x = torch.zeros(n, side, side, device=self.device)
            out = torch.zeros(n, plan.C, plan.oh, plan.ow, device=self.device)
            ws = torch.empty(chunk * plan.ws_bytes_per_image, dtype=torch.uint8, device=self.device)
            bufs.append({'side': side, 'n': n, 'plan': plan, 'filt': filt, 'x': x, 'out': out, 'ws': ws})
        k = len(bufs)
        P = ctypes.c_void_p
        arr = lambda vals: (P * k)(*[P(v) for v in vals])
        plans = arr([b['plan']._h.value for b in bufs])
        xs = arr([b['x'].data_ptr() for b in bufs])
        fs = arr([b['filt'].data_ptr() for b in bufs])
        outs = arr([b['out'].data_ptr() for b in bufs])
        wss = arr([b['ws'].data_ptr() for b in bufs])
        batch = (ctypes.c_longlong * k)(*[b['n'] for b in bufs])
        wsb = (ctypes.c_longlong * k)(*[b['ws'].numel() for b in bufs])
        self._counts = torch.zeros(k, dtype=torch.int32, device=self.device)
        g = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(lib.kym_scat2d_graph_create_counted(plans, k, xs, fs, outs, batch, wss, wsb,
                                                           P(self._counts.data_ptr()), ctypes.byref(g)))
        self._state = ({b['side']: b['n'] for b in bufs}, g, bufs)
        self.builds += 1

    def forward_grouped(self, groups):
        """groups: {side: (n, side, side) float32 tensor on the device} ->
        {side: (n, C, oh, ow)} in one graph launch."""
        counts = {}
        for side, x in groups.items():
            if not torch.is_tensor(x) or x.dim() != 3 or tuple(x.shape[1:]) != (side, side):
                raise RuntimeError('Tensor must be of spatial size (%i,%i).' % (side, side))
            if x.dtype is not torch.float32:
                raise TypeError('Input and filter must be of the same dtype.')
            _lib.require_device(x)
            if x.shape[0]:
                counts[side] = x.shape[0]
        if not counts:
            return {side: None for side in groups}
        maxima = dict(self._state[0]) if self._state is not None else {}
        if any(s not in maxima or n > maxima[s] for s, n in counts.items()):
            for s, n in counts.items():
                if s not in maxima or n > maxima[s]:
                    maxima[s] = self._headroom(n)
            self._build(maxima)
        _, g, bufs = self._state
        live = []
        for b in bufs:
            n = counts.get(b['side'], 0)
            live.append(n)
            if n:
                b['x'][:n].copy_(groups[b['side']])
        self._counts.copy_(torch.tensor(live, dtype=torch.int32), non_blocking=False)
        _lib.check(_lib.lib().kym_scat2d_graph_launch(g, _lib.stream_of(self.device)))
        return {b['side']: b['out'][:n].clone() for b, n in zip(bufs, live) if n}

    def __call__(self, images):
        """images: list of square (side, side) tensors (sides may differ) ->
        list of (C_side, oh, ow) scattering outputs in the same order."""
        by_side = {}
        for i, x in enumerate(images):
            if not torch.is_tensor(x) or x.dim() != 2 or x.shape[0] != x.shape[1]:
                raise RuntimeError('MixedScattering2D takes square 2D images.')
            by_side.setdefault(int(x.shape[0]), []).append(i)
        groups = {s: torch.stack([images[i] for i in idx]).contiguous() for s, idx in by_side.items()}
        res = self.forward_grouped(groups)
        out = [None] * len(images)
        for s, idx in by_side.items():
            for j, i in enumerate(idx):
                out[i] = res[s][j]
        return out
