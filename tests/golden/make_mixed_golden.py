"""Writes tests/golden/golden_mixed2d.npz: oracle Scattering2D outputs for four
images of side 256 (J=5) and four of side 512 (J=6), the two large sizes of the
mixed-size workload, whose CPU oracle is too slow to run inside the GPU tests
(the oracle is pinned against the reference's own vectors at other sizes).
Inputs are regenerated from the seeds: numpy default_rng(seed).random((4, s, s), float32).

    python tests/golden/make_mixed_golden.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import scat2d  # noqa: E402

CASES = [(256, 5, 11), (512, 6, 12)]


def main():
    out = {}
    for side, J, seed in CASES:
        x = np.random.default_rng(seed).random((4, side, side), dtype=np.float32)
        with torch.no_grad():
            S = scat2d.Scattering2DTorchCPU(J, (side, side))(torch.from_numpy(x)).numpy()
        out['s%d_J' % side] = np.int64(J)
        out['s%d_seed' % side] = np.int64(seed)
        out['s%d_Sx' % side] = S.astype(np.float32)
        print(side, J, S.shape, flush=True)
    np.savez_compressed(os.path.join(HERE, 'golden_mixed2d.npz'), **out)


if __name__ == '__main__':
    main()
