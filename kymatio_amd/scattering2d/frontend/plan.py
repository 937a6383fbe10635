"""Python handle of a Scattering2D plan (kym_scat2d_* in include/kymatio_hip.h)."""
import ctypes

import torch

from ... import _lib


class Scat2DPlanHandle:
    def __init__(self, M, N, J, L, max_order, pre_pad, dtype, lds_limit=0, tune=None):
        """tune: None or up to 7 ints {largest power-of-two radix, LDS-node
        threads, column block, row block, pass threads, streams, pair} (0 =
        heuristic), see kym_scat2d_create_tuned."""
        lib = _lib.lib()
        h = ctypes.c_void_p()
        self.tune = list(tune) if tune else []
        t = (ctypes.c_int * max(len(self.tune), 1))(*self.tune)
        _lib.check(lib.kym_scat2d_create_tuned(M, N, J, L, max_order, pre_pad, _lib.DTYPE_CODE[dtype],
                                               int(lds_limit), t, len(self.tune), ctypes.byref(h)))
        self._h = h
        self._lib = lib
        info = (ctypes.c_longlong * 12)()
        _lib.check(lib.kym_scat2d_info(h, info))
        (self.C, self.oh, self.ow, self.Mp, self.Np, self.n_filters, self.filter_elems,
         self.ws_bytes_per_image, self.n_launches, self.top, self.left, self.train_ws_bytes_per_image) = list(info)
        offs = (ctypes.c_longlong * (self.n_filters + 1))()
        _lib.check(lib.kym_scat2d_filter_offsets(h, offs))
        self.filter_offsets = list(offs)
        J = int(J)
        desc = (ctypes.c_longlong * (4 + 10 * max(J, 1)))()
        _lib.check(lib.kym_scat2d_leaf_taps(h, desc, len(desc)))
        self.leaf_taps_desc = list(desc)[:4 + 10 * J]
        self.dtype = dtype

    KIND_NAMES = ['lds_node', 'leaf', 'root_rows', 'icol', 'rows_mod', 'fwd_cols', 'leaf_icol', 'leaf_rows',
                  'fir_cols', 'icol_pair']

    def launches(self):
        """Per-launch records: kind, nodes, tasks per node, threads, LDS bytes,
        model bytes per image, resolution."""
        out = []
        for i in range(self.n_launches):
            rec = (ctypes.c_longlong * 7)()
            _lib.check(self._lib.kym_scat2d_launch_info(self._h, i, rec))
            r = list(rec)
            out.append({'kind': self.KIND_NAMES[r[0]], 'nodes': r[1], 'tasks_per_node': r[2], 'threads': r[3],
                        'lds_bytes': r[4], 'model_bytes_per_image': r[5], 'res': r[6]})
        return out

    def profile(self, enable):
        _lib.check(self._lib.kym_scat2d_profile(self._h, int(bool(enable))))

    def profile_read(self):
        n = self.n_launches
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        _lib.check(self._lib.kym_scat2d_profile_read(self._h, ms, cnt, n))
        return list(ms), list(cnt)

    def forward(self, x, filt, out, batch, ws):
        _lib.check(self._lib.kym_scat2d_forward(
            self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(filt.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), int(batch), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
            _lib.stream_of(x.device)))

    def forward_train(self, x, filt, out, batch, ws):
        _lib.check(self._lib.kym_scat2d_forward_train(
            self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(filt.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), int(batch), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
            _lib.stream_of(x.device)))

    def backward(self, dS, filt, gx, batch, ws):
        _lib.check(self._lib.kym_scat2d_backward(
            self._h, ctypes.c_void_p(dS.data_ptr()), ctypes.c_void_p(filt.data_ptr()),
            ctypes.c_void_p(gx.data_ptr()), int(batch), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
            _lib.stream_of(dS.device)))

    def __del__(self):
        try:
            if self._h:
                self._lib.kym_scat2d_destroy(self._h)
                self._h = None
        except Exception:
            pass
