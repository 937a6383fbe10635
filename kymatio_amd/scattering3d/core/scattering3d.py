"""3D solid harmonic scattering over a backend namespace (restates
kymatio/scattering3d/core/scattering3d.py:1-75): for each l and j1 the
(rotation-covariant) modulus of x * psi_{l, j1, m}, its integrals, and at
order 2 the same from |x * psi_{j1}| with j2 > j1; outputs ordered (j
paths, l)."""


def _moduli(U_hat, filters_lj, rotation_covariant, backend):
    if rotation_covariant:
        U = None
        for f in filters_lj:
            U = backend.modulus_rotation(backend.ifft(backend.cdgmm3d(U_hat, f)), U)
        return U
    return backend.modulus(backend.ifft(backend.cdgmm3d(U_hat, filters_lj[0])))


def scattering3d(x, filters, rotation_covariant, L, J, max_order, backend, averaging):
    U_0_hat = backend.rfft(x)
    per_l = []
    for l in range(L + 1):
        s1, s2 = [], []
        for j1 in range(J + 1):
            U_1 = _moduli(U_0_hat, filters[l][j1], rotation_covariant, backend)
            s1.append(averaging(U_1))
            if max_order > 1:
                U_1_hat = backend.rfft(U_1)
                for j2 in range(j1 + 1, J + 1):
                    s2.append(averaging(_moduli(U_1_hat, filters[l][j2], rotation_covariant, backend)))
        per_l.append(s1 + s2 if max_order == 2 else s1)
    # (l, paths) -> (paths, l)
    S = [p for group in zip(*per_l) for p in group]
    return backend.stack(S, L)
