"""Joint time-frequency scattering (JTFS) over a backend namespace of ops.

Restates the algorithm of kymatio/scattering1d/core/timefrequency_scattering.py
(joint_timefrequency_scattering :4-81, time_scattering_widthfirst :84-189,
frequency_scattering :192-275, time_averaging :278-315, frequency_averaging
:318-373, time_formatting :376-392, jtfs_average_and_format :395-444): the
same paths, in the same order, with the same keys, so outputs match the
reference op for op.  With the HIP backend every transform, product,
periodisation and modulus below is one HIP kernel (kymatio_amd.scattering1d.
backend); the frequency-axis ops are tensor layout moves.

Tensors are (batch, ..., time, 1 | 2) like the reference's torch backend;
the log-frequency axis n1 is the one before time.
"""
import numpy as np


def _time_filter(backend, x_hat, filt, level, k):
    """cdgmm + periodise by 2^k (the spectrum of x * filt subsampled)."""
    return backend.subsample_fourier(backend.cdgmm(x_hat, filt['levels'][level]), 2 ** k)


def _filtered(backend, x_hat, filt, level, k, mode):
    """ifft (mode 'complex'), irfft ('real') or modulus of ifft ('modulus') of
    _time_filter(x_hat, filt, level, k): one fused op where the backend has
    it (the HIP backend's filter_ifft), else the reference's op sequence."""
    fused = getattr(backend, 'filter_ifft', None)
    if fused is not None:
        return fused(x_hat, filt['levels'][level], 2 ** k, mode)
    y = _time_filter(backend, x_hat, filt, level, k)
    if mode == 'real':
        return backend.irfft(y)
    y = backend.ifft(y)
    return backend.modulus(y) if mode == 'modulus' else y


def psi1_runs(psi1, log2_stride, average_local):
    """Runs [a, b) of consecutive n1 with one j1 (so one k1): (a, b, j1, k1)."""
    runs = []
    for n1, p1 in enumerate(psi1):
        j1 = p1['j']
        k1 = min(j1, log2_stride) if average_local else j1
        if runs and runs[-1][2] == j1:
            runs[-1] = (runs[-1][0], n1 + 1, j1, k1)
        else:
            runs.append((n1, n1 + 1, j1, k1))
    return runs


def time_scattering_widthfirst(U_0, backend, filters, log2_stride, average_local, banks=None):
    """Zeroth order, then S_1 stacked over n1 (batch, n1, time), then for each
    n2 the complex Y_2 stacked over the n1 with j1 < j2 (width first).
    banks ({a: the psi1 filters of run a, stacked}, psi1_runs) with a
    backend that has filter_ifft_bank: every op runs once per run of n1 with
    one j1 instead of once per n1, the same values in the same layout."""
    phi, psi1, psi2 = filters
    U_0_hat = backend.rfft(U_0)
    if average_local:
        yield {'coef': _filtered(backend, U_0_hat, phi, 0, log2_stride, 'real'), 'j': (), 'n': ()}
    else:
        yield {'coef': U_0, 'j': (), 'n': ()}
    if banks is not None and getattr(backend, 'filter_ifft_bank', None) is not None:
        yield from _widthfirst_by_runs(U_0_hat, backend, filters, log2_stride, average_local, banks)
        return

    U_1_hats, S_1 = [], []
    for n1, p1 in enumerate(psi1):
        j1 = p1['j']
        k1 = min(j1, log2_stride) if average_local else j1
        U_1 = _filtered(backend, U_0_hat, p1, 0, k1, 'modulus')
        U_1_hat = backend.rfft(U_1)
        U_1_hats.append((j1, k1, U_1_hat))
        S_1.append(_filtered(backend, U_1_hat, phi, k1, log2_stride - k1, 'real'))
    yield {'coef': backend.stack(S_1, 2), 'j': (-1,), 'n': (-1,), 'n1_max': len(S_1)}

    for n2, p2 in enumerate(psi2):
        j2 = p2['j']
        Y_2 = []
        for j1, k1, U_1_hat in U_1_hats:
            if j2 > j1:
                k2 = min(j2 - k1, log2_stride) if average_local else j2 - k1
                Y_2.append(_filtered(backend, U_1_hat, p2, k1, k2, 'complex'))
        if Y_2:
            yield {'coef': backend.stack(Y_2, 2), 'j': (-1, j2), 'n': (-1, n2), 'n1_max': len(Y_2)}


def _widthfirst_by_runs(U_0_hat, backend, filters, log2_stride, average_local, banks):
    phi, psi1, psi2 = filters
    runs = psi1_runs(psi1, log2_stride, average_local)
    n1_max = len(psi1)
    U_1_hats, S_1 = [], None
    for a, b, j1, k1 in runs:
        U_1 = backend.filter_ifft_bank(U_0_hat, banks[a], 2 ** k1, 'modulus')
        U_1_hat = backend.rfft(U_1)
        U_1_hats.append((a, b, j1, k1, U_1_hat))
        S_1 = backend.filter_ifft_into(U_1_hat, phi['levels'][k1], 2 ** (log2_stride - k1), 'real', S_1, a, n1_max)
    yield {'coef': S_1, 'j': (-1,), 'n': (-1,), 'n1_max': n1_max}

    for n2, p2 in enumerate(psi2):
        j2 = p2['j']
        members = [r for r in U_1_hats if j2 > r[2]]
        count = sum(r[1] - r[0] for r in members)
        if not count:
            continue
        Y_2, off = None, 0
        for a, b, j1, k1, U_1_hat in members:
            k2 = min(j2 - k1, log2_stride) if average_local else j2 - k1
            Y_2 = backend.filter_ifft_into(U_1_hat, p2['levels'][k1], 2 ** k2, 'complex', Y_2, off, count)
            off += b - a
        yield {'coef': Y_2, 'j': (-1, j2), 'n': (-1, n2), 'n1_max': count}


def frequency_scattering(X, backend, filters_fr, log2_stride_fr, average_local_fr, spinned, modulus=False):
    """Wavelets along n1 (zero-padded to the frequency filters' length N):
    all n_fr for complex input (spinned), only spin >= 0 for real input.
    modulus: yield |Y| (the modulus jtfs_average_and_format takes next):
    every wavelet from one read of X where the backend has freq_scatter,
    else filter, inverse, swap and modulus as one op per wavelet where it
    has filter_ifft_swap."""
    phi, psis = filters_fr
    fused = getattr(backend, 'freq_scatter', None) if modulus else None
    if fused is not None:  # every wavelet in one pass over X (freq_scatter)
        if not spinned:
            psis = [p for p in psis if p['xi'] >= 0]
        ks = [2 ** (min(p['j'], log2_stride_fr) if average_local_fr else p['j']) for p in psis]
        Ys = fused(X['coef'], [p['levels'][0] for p in psis], ks, phi['N'])
        for n_fr, (psi, Y) in enumerate(zip(psis, Ys)):
            j_fr = psi['j']
            yield {**X, 'coef': Y, 'n': X['n'][1:] + (n_fr,), 'j_fr': (j_fr,), 'n_fr': (n_fr,),
                   'n1_stride': 2 ** j_fr, 'spin': np.sign(psi['xi'])}
        return
    X_T = backend.pad_frequency(backend.swap_time_frequency(X['coef']), phi['N'] - X['n1_max'])
    if spinned:
        X_hat = backend.cfft(X_T)
    else:
        X_hat = backend.rfft(X_T)
        psis = [p for p in psis if p['xi'] >= 0]
    for n_fr, psi in enumerate(psis):
        j_fr = psi['j']
        k_fr = min(j_fr, log2_stride_fr) if average_local_fr else j_fr
        fused = getattr(backend, 'filter_ifft_swap', None) if modulus else None
        if fused is not None:
            Y = fused(X_hat, psi['levels'][0], 2 ** k_fr, 'modulus')
        else:
            Y = backend.swap_time_frequency(_filtered(backend, X_hat, psi, 0, k_fr, 'complex'))
            if modulus:
                Y = backend.modulus(Y)
        yield {**X, 'coef': Y, 'n': X['n'][1:] + (n_fr,), 'j_fr': (j_fr,), 'n_fr': (n_fr,),
               'n1_stride': 2 ** j_fr, 'spin': np.sign(psi['xi'])}


def joint_timefrequency_scattering(U_0, backend, filters, log2_stride, average_local, filters_fr, log2_stride_fr,
                                   average_local_fr, modulus=False, banks=None):
    """Zeroth order; frequency scattering of S_1 (real, spin >= 0); then for
    each n2 frequency scattering of Y_2 (complex, both spins).  modulus:
    paths after the first come as their modulus (frequency_scattering);
    banks: see time_scattering_widthfirst."""
    time_gen = time_scattering_widthfirst(U_0, backend, filters, log2_stride, average_local, banks)
    yield next(time_gen)
    S_1 = next(time_gen)
    yield from frequency_scattering(S_1, backend, filters_fr, log2_stride_fr, average_local_fr, spinned=False,
                                    modulus=modulus)
    for Y_2 in time_gen:
        yield from frequency_scattering(Y_2, backend, filters_fr, log2_stride_fr, average_local_fr, spinned=True,
                                        modulus=modulus)


def time_averaging(U_2, backend, phi_f, log2_stride):
    """Low-pass phi_f along time, subsampled to 2^log2_stride: rfft then the
    real filtered inverse, or one op where the backend has it (lowpass)."""
    k_in = U_2['j'][-1]
    lowpass = getattr(backend, 'lowpass', None)
    if lowpass is not None:
        return {**U_2, 'coef': lowpass(U_2['coef'], phi_f['levels'][k_in], 2 ** (log2_stride - k_in), -2)}
    U_hat = backend.rfft(U_2['coef'])
    return {**U_2, 'coef': _filtered(backend, U_hat, phi_f, k_in, log2_stride - k_in, 'real')}


def frequency_averaging(U_2, backend, phi_fr_f, log2_stride_fr, average_fr):
    if not average_fr:
        return {**U_2, 'n1_stride': 2 ** max(U_2['j_fr'][-1], 0)}
    if average_fr == 'global':
        S_T = backend.average_global(backend.swap_time_frequency(U_2['coef']))
        n1_stride = U_2['n1_max']
    else:  # 'local'
        k_in = min(U_2['j_fr'][-1], phi_fr_f['j'])
        n1_stride = 2 ** log2_stride_fr
        lowpass = getattr(backend, 'lowpass', None)
        if lowpass is not None:  # along n1 in place of the two swaps
            S = lowpass(U_2['coef'], phi_fr_f['levels'][k_in], 2 ** (log2_stride_fr - k_in), -3)
            return {**U_2, 'coef': S, 'n1_stride': n1_stride}
        U_T = backend.swap_time_frequency(U_2['coef'])
        S_T = _filtered(backend, backend.rfft(U_T), phi_fr_f, k_in, log2_stride_fr - k_in, 'real')
    return {**U_2, 'coef': backend.swap_time_frequency(S_T), 'n1_stride': n1_stride}


def time_formatting(path, backend):
    """One path per kept n1 (n1 = 0, n1_stride, ...); n gains n1 in front."""
    n1s = range(0, path['n1_max'], path['n1_stride'])
    coefs = [None] * (1 + path['n1_max'] // path['n1_stride']) if path['coef'] is None else \
        backend.split_frequency_axis(path['coef'])
    for i, n1 in enumerate(n1s):
        out = {k: v for k, v in path.items() if k not in ('n1_max', 'n1_stride')}
        out.update(coef=coefs[i], order=len(path['n']) - 1, n=(n1,) + path['n'])
        yield out


def jtfs_average_and_format(U_gen, backend, phi_f, log2_stride, average, phi_fr_f, log2_stride_fr, average_fr,
                            out_type, format, modulus_done=False):
    """modulus_done: U_gen's paths after the first are already moduli
    (joint_timefrequency_scattering with modulus=True)."""
    path = next(U_gen)
    if average == 'global':
        path['coef'] = backend.average_global(path['coef'])
    yield {**path, 'order': 0}
    for path in U_gen:
        if not modulus_done:
            path['coef'] = backend.modulus(path['coef'])
        if average == 'global':
            path['coef'] = backend.average_global(path['coef'])
        elif average == 'local' and len(path['n']) > 1:
            path = time_averaging(path, backend, phi_f, log2_stride)
        # spin 0 (the low-pass along n1) is already averaged in frequency
        if average_fr and path['spin'] != 0:
            path = frequency_averaging(path, backend, phi_fr_f, log2_stride_fr, average_fr)
        if not (out_type == 'array' and format == 'joint'):
            path['coef'] = backend.unpad_frequency(path['coef'], path['n1_max'], path['n1_stride'])
        if format == 'joint':
            yield {**path, 'order': len(path['n'])}
        else:
            yield from time_formatting(path, backend)
