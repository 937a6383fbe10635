"""Write golden vectors for the parts of the path the reference's own fixtures
do not pin, by running the reference (kymatio 0.4.0.dev0, pure Python) on the
CPU from /root/reference.

Run from the repository root in the build container (the reference does not
exist on the GPU machines):

    python tests/golden/make_golden.py

Outputs (data only: inputs, seeds and expected outputs):
    tests/golden/golden_filters2d.npz  2D filter banks at P = 40, 160, 272
                                       (full at P=40; per-filter sum, L2 norm and
                                       64 sampled entries otherwise)
    tests/golden/golden_scat2d.npz     torch-backend S(x) at the workload shapes,
                                       odd / non-square sizes, pre_pad, max_order=1,
                                       and d/dx of sum((S(x)-S(y))^2) at J=4 128x128
    tests/golden/golden_scat1d.npz     1D filters at N_pad=8192 J=6 Q=8 (checksums),
                                       S(x) for the chirp workload and T/stride variants
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, '/root/reference')

from kymatio.scattering2d.filter_bank import filter_bank  # noqa: E402
from kymatio.torch import Scattering1D, Scattering2D  # noqa: E402

from kymatio_amd import workloads  # noqa: E402


def summarize(arr, rng, nsamp=64):
    flat = arr.reshape(-1)
    idx = rng.integers(0, flat.size, size=nsamp)
    return (np.float64(flat.astype(np.float64).sum()),
            np.float64(np.sqrt((flat.astype(np.float64) ** 2).sum())), idx, flat[idx])


def filters2d():
    out = {}
    for tag, (P, J) in {'P40': (40, 2), 'P160': (160, 4), 'P272': (272, 3)}.items():
        fb = filter_bank(P, P, J, 8)
        rng = np.random.default_rng(P)
        names, sums, norms, idxs, vals, fulls = [], [], [], [], [], []
        allf = [('phi', r, fb['phi']['levels'][r]) for r in range(J)]
        for n, p in enumerate(fb['psi']):
            for r, lev in enumerate(p['levels']):
                allf.append(('psi%d' % n, r, lev))
        for name, r, lev in allf:
            s, nrm, idx, v = summarize(lev, rng)
            names.append('%s_r%d' % (name, r))
            sums.append(s)
            norms.append(nrm)
            idxs.append(idx)
            vals.append(v)
            if P == 40:
                out['%s_full_%s_r%d' % (tag, name, r)] = lev
        out[tag + '_names'] = np.array(names)
        out[tag + '_sum'] = np.array(sums)
        out[tag + '_norm'] = np.array(norms)
        out[tag + '_idx'] = np.stack(idxs)
        out[tag + '_val'] = np.stack(vals)
        out[tag + '_J'] = J
    np.savez_compressed(os.path.join(HERE, 'golden_filters2d.npz'), **out)


def scat2d():
    out = {}

    def run(tag, x, J, shape, **kw):
        S = Scattering2D(J, shape=shape, **kw)
        xt = torch.from_numpy(x)
        with torch.no_grad():
            Sx = S(xt)
        out[tag + '_x'] = x
        out[tag + '_Sx'] = Sx.numpy()
        out[tag + '_J'] = J
        out[tag + '_shape'] = np.array(shape)
        for k, v in kw.items():
            out[tag + '_' + k] = v

    run('cfg1', workloads.uniform_images(4, 32, seed=1), 2, (32, 32))
    run('cfg1o1', workloads.uniform_images(4, 32, seed=1), 2, (32, 32), max_order=1)
    run('cfg2', workloads.grating_images(1, 256, seed=2), 3, (256, 256))
    run('cfg4', workloads.normal_images(2, 128, seed=4), 4, (128, 128))
    rng = np.random.default_rng(7)
    run('odd31', rng.standard_normal((2, 31, 31)).astype(np.float32), 1, (31, 31))
    run('odd33', rng.standard_normal((2, 33, 33)).astype(np.float32), 2, (33, 33))
    run('rect37x32', rng.standard_normal((2, 37, 32)).astype(np.float32), 5, (37, 32))
    run('sq32J5', rng.standard_normal((1, 32, 32)).astype(np.float32), 5, (32, 32))
    run('rect40x24', rng.standard_normal((2, 40, 24)).astype(np.float32), 3, (40, 24))
    run('prepad', rng.standard_normal((2, 40, 40)).astype(np.float32), 2, (32, 32),
        pre_pad=True)
    run('L6', rng.standard_normal((2, 48, 48)).astype(np.float32), 3, (48, 48), L=6)

    # gradient of the config-4 loss through the reference (torch autograd)
    x = workloads.normal_images(1, 128, seed=40)
    y = workloads.normal_images(1, 128, seed=41)
    for dt, name in ((torch.float32, 'f32'), (torch.float64, 'f64')):
        S = Scattering2D(4, shape=(128, 128))
        if dt == torch.float64:
            S = S.double()
        xt = torch.from_numpy(x).to(dt).requires_grad_()
        yt = torch.from_numpy(y).to(dt)
        loss = ((S(xt) - S(yt)) ** 2).sum()
        loss.backward()
        out['grad_' + name + '_loss'] = loss.item()
        out['grad_' + name + '_dx'] = xt.grad.numpy()
    out['grad_x'] = x
    out['grad_y'] = y
    np.savez_compressed(os.path.join(HERE, 'golden_scat2d.npz'), **out)


def scat1d():
    out = {}
    S = Scattering1D(6, 4096, 8)
    rng = np.random.default_rng(8192)
    names, sums, norms, idxs, vals = [], [], [], [], []
    banks = [('phi', [S.phi_f])] + [('psi1_', S.psi1_f), ('psi2_', S.psi2_f)]
    for pre, bank in banks:
        for n, p in enumerate(bank):
            for r, lev in enumerate(p['levels']):
                arr = lev.numpy().reshape(-1) if torch.is_tensor(lev) else np.asarray(lev)
                s, nrm, idx, v = summarize(arr, rng)
                names.append('%s%d_r%d' % (pre, n, r))
                sums.append(s)
                norms.append(nrm)
                idxs.append(idx)
                vals.append(v)
    out['f8192_names'] = np.array(names)
    out['f8192_sum'] = np.array(sums)
    out['f8192_norm'] = np.array(norms)
    out['f8192_idx'] = np.stack(idxs)
    out['f8192_val'] = np.stack(vals)
    out['f8192_xi1'] = np.array([p['xi'] for p in S.psi1_f])
    out['f8192_sigma1'] = np.array([p['sigma'] for p in S.psi1_f])
    out['f8192_j1'] = np.array([p['j'] for p in S.psi1_f])
    out['f8192_xi2'] = np.array([p['xi'] for p in S.psi2_f])
    out['f8192_j2'] = np.array([p['j'] for p in S.psi2_f])

    def run(tag, x, J, N, Q, **kw):
        Sm = Scattering1D(J, N, Q, **kw)
        with torch.no_grad():
            Sx = Sm(torch.from_numpy(x))
        out[tag + '_x'] = x
        out[tag + '_Sx'] = Sx.numpy()
        out[tag + '_J'] = J
        out[tag + '_Q'] = Q
        for k, v in kw.items():
            out[tag + '_' + k] = v

    run('cfg5', workloads.chirp_signals(2, 4096, seed=5), 6, 4096, 8)
    rng = np.random.default_rng(11)
    x512 = rng.standard_normal((2, 512)).astype(np.float32)
    run('T16', x512, 6, 512, 16, T=16)
    run('stride4', x512, 6, 512, 16, stride=4)
    run('Q12', x512, 5, 512, (12, 2))
    run('o1', x512, 6, 512, 16, max_order=1)
    run('N1000', rng.standard_normal((2, 1000)).astype(np.float32), 5, 1000, 4)
    np.savez_compressed(os.path.join(HERE, 'golden_scat1d.npz'), **out)


if __name__ == '__main__':
    torch.manual_seed(0)
    filters2d()
    scat2d()
    scat1d()
    print('written to', HERE)
