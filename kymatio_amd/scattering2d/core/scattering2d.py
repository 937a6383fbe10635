"""Per-op 2D scattering loop over a backend namespace, for user-supplied
backend objects (the plugin route of kymatio/frontend/base_frontend.py:32-50).
Same path order and op sequence as kymatio/scattering2d/core/scattering2d.py:1-88."""


def scattering2d(x, pad, unpad, backend, J, L, phi, psi, max_order, out_type='array'):
    B = backend
    U0 = B.rfft(pad(x))
    s0 = unpad(B.irfft(B.subsample_fourier(B.cdgmm(U0, phi['levels'][0]), k=2 ** J)))
    first, second = [], []
    for n1, p1 in enumerate(psi):
        j1 = p1['j']
        U1 = B.cdgmm(U0, p1['levels'][0])
        if j1 > 0:
            U1 = B.subsample_fourier(U1, k=2 ** j1)
        U1 = B.rfft(B.modulus(B.ifft(U1)))
        S1 = B.subsample_fourier(B.cdgmm(U1, phi['levels'][j1]), k=2 ** (J - j1))
        first.append({'coef': unpad(B.irfft(S1)), 'j': (j1,), 'n': (n1,), 'theta': (p1['theta'],)})
        if max_order < 2:
            continue
        for n2, p2 in enumerate(psi):
            j2 = p2['j']
            if j2 <= j1:
                continue
            U2 = B.subsample_fourier(B.cdgmm(U1, p2['levels'][j1]), k=2 ** (j2 - j1))
            U2 = B.rfft(B.modulus(B.ifft(U2)))
            S2 = B.subsample_fourier(B.cdgmm(U2, phi['levels'][j2]), k=2 ** (J - j2))
            second.append({'coef': unpad(B.irfft(S2)), 'j': (j1, j2), 'n': (n1, n2),
                           'theta': (p1['theta'], p2['theta'])})
    out = [{'coef': s0, 'j': (), 'n': (), 'theta': ()}] + first + second
    if out_type == 'array':
        return B.stack([o['coef'] for o in out])
    return out
