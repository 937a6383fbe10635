"""What the trained decoder gains: wall time for Adam refinement through the
HIP Scattering1D backward to bring the relative scattering loss
L = ||S(y) - S(x)||^2 / ||S(x)||^2 of a batch to `tol`, started from the
decoder's output (learned warm start) or from N(0, 1) noise (plain descent).
The decoder's forward is counted in the warm-start time.  Each start uses
the best of a few Adam step sizes (the same list for both).

    python -m kymatio_amd.learned.evaluate [--checkpoint PATH] [--tol 1e-3]
"""
import argparse
import json
import time

import torch

from .decoder import ScatteringDecoder1D
from .recon import reconstruct, rel_loss
from .train import CHECKPOINT, chirps

LRS = (0.3, 0.1, 0.03)


def load_decoder(path, device):
    state = torch.load(path, map_location='cpu', weights_only=True)
    C, T_in, n_out, width = [int(v) for v in state['config']]
    dec = ScatteringDecoder1D(C, T_in, n_out, width=width, device=device)
    dec.load_state_dict(state)
    return dec


def _run(scattering, S, y0, tol, lr, time_limit, max_iter, extra_s=0.0):
    y, hist, ok = reconstruct(scattering, S, y0, tol=tol, max_iter=max_iter, lr=lr, check_every=5,
                              time_limit=time_limit)
    it, t, L = hist[-1]
    return {'lr': lr, 'reached': bool(ok), 'iterations': it, 'seconds': t + extra_s, 'final_loss': L}


def compare(scattering, decoder, n=256, seed=99, tol=1e-3, lrs=LRS, time_limit=30.0, max_iter=20000):
    dev = next(scattering.buffers()).device
    x = chirps(n, scattering.shape[0], seed, 0, device=dev)
    with torch.no_grad():
        S = scattering(x)
    g = torch.Generator(device=dev).manual_seed(seed)
    noise = torch.randn(x.shape, device=dev, generator=g)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    with torch.no_grad():
        y0 = decoder.forward(S)
    torch.cuda.synchronize(dev)
    t_dec = time.perf_counter() - t0
    out = {'tol': tol, 'n_signals': n, 'loss_definition': '||S(y)-S(x)||^2 / ||S(x)||^2 over the batch',
           'start_loss': {'plain': rel_loss(scattering, noise, S), 'warm': rel_loss(scattering, y0, S)},
           'decoder_forward_s': t_dec}
    for name, start, extra in (('plain', noise, 0.0), ('warm', y0, t_dec)):
        runs = [_run(scattering, S, start, tol, lr, time_limit, max_iter, extra) for lr in lrs]
        done = [r for r in runs if r['reached']]
        out[name] = min(done, key=lambda r: r['seconds']) if done else min(runs, key=lambda r: r['final_loss'])
        out[name + '_runs'] = runs
    if out['plain']['reached'] and out['warm']['reached']:
        out['speedup'] = out['plain']['seconds'] / out['warm']['seconds']
    else:
        out['speedup'] = None
    return out


def main(argv=None):
    from ..torch import Scattering1D
    ap = argparse.ArgumentParser()
    ap.add_argument('--checkpoint', default=CHECKPOINT)
    ap.add_argument('--tol', type=float, default=1e-3)
    ap.add_argument('--time-limit', type=float, default=30.0)
    args = ap.parse_args(argv)
    dev = torch.device('cuda')
    sc = Scattering1D(6, 4096, 8).to(dev)
    dec = load_decoder(args.checkpoint, dev)
    print(json.dumps(compare(sc, dec, tol=args.tol, time_limit=args.time_limit)))


if __name__ == '__main__':
    main()
