"""Solid harmonic scattering as batched HIP launches.

The reference walks l, j1 (and j2 > j1) and, inside, every m of the order-l
wavelets one op at a time: cdgmm3d, ifft, modulus_rotation per m, then one
integral per path (kymatio/scattering3d/core/scattering3d.py:24-75).  Here
each (l, j) step is three launches over all m and all images together:

    products  (B, 2l+1, n) = U_hat (B, n) x psi_{l, j, m} (2l+1, n)   kym_op_harmonic_products
    inverse   3D FFT of the B (2l+1) volumes                          backend.ifft
    modulus   (B, n) = the chained rotation modulus over m            kym_op_modulus_rotation_sum

and the moduli of every (path, l) land in one field array of layout
(paths, L + 1, B, n), so one kym_op_integrals launch computes every output
and the result is that array's integrals moved to (B, paths, L + 1, P).
Path order per l is the reference's: all first-order j1, then the pairs
(j1, j2 > j1) in j1-major order; outputs interleave l inside each path
(scattering3d.py:64-73).
"""
import ctypes

import torch

from ... import _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def harmonic_paths(J, max_order):
    """Paths in output order: (j1,) for j1 <= J, then (j1, j2), j1 < j2 <= J."""
    first = [(j1,) for j1 in range(J + 1)]
    if max_order < 2:
        return first
    return first + [(j1, j2) for j1 in range(J + 1) for j2 in range(j1 + 1, J + 1)]


class _Launcher:
    """The batched ops on one device and stream, for one dtype."""

    def __init__(self, device, dtype):
        self.lib = _lib.lib()
        self.device = device
        self.stream = _lib.stream_of(device)
        self.code = _lib.DTYPE_CODE[dtype]

    def products(self, U_hat, F):
        B, nm = U_hat.shape[0], F.shape[0]
        n = U_hat[0].numel() // 2
        out = torch.empty((B, nm) + tuple(U_hat.shape[1:]), dtype=U_hat.dtype, device=self.device)
        _lib.check(self.lib.kym_op_harmonic_products(_ptr(U_hat), _ptr(F), _ptr(out), B, nm, n, self.code,
                                                       self.stream))
        return out

    def rotation_modulus(self, Y, dest):
        """Y: (B, nm, M, N, O, 2) -> dest (B, n) (a contiguous block of the field array)."""
        B, nm = Y.shape[0], Y.shape[1]
        n = dest.shape[1]
        _lib.check(self.lib.kym_op_modulus_rotation_sum(_ptr(Y), _ptr(dest), B, nm, n, n, self.code, self.stream))

    def integrals(self, fields, powers):
        rows, n = fields.shape[0], fields.shape[1]
        P = len(powers)
        pw = (ctypes.c_double * P)(*[float(p) for p in powers])
        out = torch.empty((rows, P), dtype=fields.dtype, device=self.device)
        _lib.check(self.lib.kym_op_integrals(_ptr(fields), _ptr(out), rows, n, pw, P, self.code, self.stream))
        return out


def harmonic_scattering(x, bank, J, L, max_order, rotation_covariant, powers, backend):
    """x: (B, M, N, O, 1) real, contiguous, on the device; bank[l]: the
    (J + 1, 2l + 1, M, N, O, 2) wavelets of order l.  Returns
    (B, paths, L + 1, len(powers)) in float32 (the dtype the reference's
    integrals are allocated in, kymatio/scattering3d/backend/torch_backend.py:145-150)."""
    B = x.shape[0]
    vol = tuple(x.shape[1:4])
    n = vol[0] * vol[1] * vol[2]
    paths = harmonic_paths(J, max_order)
    slot = {p: i for i, p in enumerate(paths)}
    run = _Launcher(x.device, x.dtype)
    fields = torch.empty((len(paths), L + 1, B, n), dtype=x.dtype, device=x.device)

    def modulus_of(U_hat, l, j, dest):
        F = bank[l][j] if rotation_covariant else bank[l][j][:1]
        run.rotation_modulus(backend.ifft(run.products(U_hat, F).reshape((-1,) + vol + (2,))).reshape(
            (B, F.shape[0]) + vol + (2,)), dest)

    U0_hat = backend.rfft(x)
    for l in range(L + 1):
        for j1 in range(J + 1):
            first = fields[slot[(j1,)], l]
            modulus_of(U0_hat, l, j1, first)
            if max_order > 1 and j1 < J:
                U1_hat = backend.rfft(first.reshape((B,) + vol + (1,)))
                for j2 in range(j1 + 1, J + 1):
                    modulus_of(U1_hat, l, j2, fields[slot[(j1, j2)], l])
    S = run.integrals(fields.reshape(-1, n), powers).reshape(len(paths), L + 1, B, len(powers))
    return S.permute(2, 0, 1, 3).contiguous().to(torch.float32)


__all__ = ['harmonic_paths', 'harmonic_scattering']
