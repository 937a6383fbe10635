"""Write golden vectors for time-frequency scattering (JTFS), which neither
the reference's own fixtures nor a closed form pin: the reference (kymatio
0.4.0.dev0, pure Python, torch frontend on the CPU in float32) from
/root/reference on seeded chirp signals.

Run from the repository root in the build container (the reference does not
exist on the GPU machines):

    python tests/golden/make_jtfs_golden.py

Output (data only): tests/golden/golden_jtfs.npz with, per case, the input,
the constructor arguments (JSON), the output (arrays: as returned; dict /
list: the coefficients flattened and concatenated in the output's order,
plus the path keys), the frequency filters' first levels, and meta()'s keys.
"""
import json
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, '/root/reference')

from kymatio.torch import TimeFrequencyScattering  # noqa: E402  (the reference)
from kymatio_amd import workloads  # noqa: E402

CASES = [
    ('time', dict(J=6, J_fr=2, Q=8, shape=1024), 2, 11),
    ('joint', dict(J=6, J_fr=2, Q=8, shape=1024, format='joint'), 2, 12),
    ('dictF0', dict(J=5, J_fr=3, Q=4, shape=512, F=0, out_type='dict'), 2, 13),
    ('listglobal', dict(J=5, J_fr=2, Q=8, shape=512, F='global', out_type='list'), 2, 14),
    ('stride', dict(J=5, J_fr=3, Q=(8, 2), shape=1000, F=4, stride_fr=2), 2, 15),
]


def main():
    out = {}
    for tag, kw, batch, seed in CASES:
        x = workloads.chirp_signals(batch, kw['shape'], seed=seed)
        S = TimeFrequencyScattering(**kw)
        with torch.no_grad():
            Sx = S(torch.from_numpy(x))
        out[tag + '_x'] = x
        out[tag + '_kw'] = np.array(json.dumps(kw))
        if isinstance(Sx, dict):
            keys = list(Sx.keys())
            out[tag + '_keys'] = np.array(json.dumps([list(k) for k in keys]))
            out[tag + '_flat'] = np.concatenate([Sx[k].numpy().reshape(batch, -1) for k in keys], axis=1)
        elif isinstance(Sx, list):
            out[tag + '_keys'] = np.array(json.dumps([list(p['n']) for p in Sx]))
            out[tag + '_flat'] = np.concatenate([p['coef'].numpy().reshape(batch, -1) for p in Sx], axis=1)
        else:
            out[tag + '_Sx'] = Sx.numpy()
        phi_fr, psis_fr = S.filters_fr
        out[tag + '_phi_fr'] = np.asarray(phi_fr['levels'][0]).reshape(-1)
        out[tag + '_psi_fr'] = np.stack([np.asarray(p['levels'][0]).reshape(-1) for p in psis_fr])
        out[tag + '_psi_fr_xi'] = np.array([p['xi'] for p in psis_fr])
        out[tag + '_meta_keys'] = np.array(json.dumps([list(k) for k in S.meta()['key']]))
        print(tag, {k: v.shape for k, v in out.items() if k.startswith(tag + '_')}, flush=True)
    np.savez_compressed(os.path.join(HERE, 'golden_jtfs.npz'), **out)


if __name__ == '__main__':
    main()
