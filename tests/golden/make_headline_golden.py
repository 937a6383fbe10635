"""Write reference outputs for more images of the headline workload
(Scattering2D J=3 L=8 on 256x256 float32 gratings + N(0,1) noise) by running
the reference (kymatio 0.4.0.dev0, pure Python, torch backend, CPU) from
/root/reference.  Run from the repository root in the build container:

    python tests/golden/make_headline_golden.py

Output: tests/golden/golden_headline2d.npz with x (3, 256, 256) float32
(kymatio_amd.workloads.grating_images(3, 256, seed=12)) and Sx
(3, 217, 32, 32) float32, the reference's torch-frontend output.
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, '/root/reference')

from kymatio.torch import Scattering2D  # noqa: E402

from kymatio_amd import workloads  # noqa: E402

if __name__ == '__main__':
    torch.manual_seed(0)
    x = workloads.grating_images(3, 256, seed=12)
    S = Scattering2D(3, shape=(256, 256))
    with torch.no_grad():
        Sx = S(torch.from_numpy(x)).numpy()
    np.savez_compressed(os.path.join(HERE, 'golden_headline2d.npz'), x=x, Sx=Sx, J=3, seed=12)
    print('written', Sx.shape)
