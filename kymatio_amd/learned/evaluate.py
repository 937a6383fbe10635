"""What the trained decoder gains: wall time for descent through the HIP
Scattering1D backward to bring the relative scattering error of a batch,

    rel = ||S(y) - S(x)||_F / ||S(x)||_F      (un-squared; the norm of
                                              examples/1d/reconstruct_torch.py:104),

to `tol`, started from the decoder's output (learned warm start) or from
N(0, 1) noise (plain descent).  The decoder's forward is counted in the
warm-start time.  Both starts run the same method with the same settings.

    python -m kymatio_amd.learned.evaluate [--checkpoint PATH] [--tol 1e-3] [--method lbfgs]
"""
import argparse
import json
import time

import torch

from .decoder import ScatteringDecoder1D
from .recon import reconstruct, rel_error
from .train import CHECKPOINT, chirps

LOSS_DEFINITION = '||S(y)-S(x)||_F / ||S(x)||_F over the batch (un-squared)'
LEVELS = (1e-1, 5e-2, 3e-2, 2e-2, 1e-2, 5e-3, 2e-3, 1e-3)


def time_to_levels(hist, extra_s=0.0, levels=LEVELS):
    """Seconds (and iterations) at which a descent history first reached
    each relative error level; None where it did not."""
    out = {}
    for lev in levels:
        hit = next((h for h in hist if h[2] <= lev), None)
        out['%g' % lev] = None if hit is None else {'seconds': hit[1] + extra_s, 'iterations': hit[0]}
    return out


def load_decoder(path, device):
    state = torch.load(path, map_location='cpu', weights_only=True)
    C, T_in, n_out, width = [int(v) for v in state['config']]
    dec = ScatteringDecoder1D(C, T_in, n_out, width=width, device=device)
    dec.load_state_dict(state)
    return dec


def _run(scattering, S, y0, tol, method, lr, time_limit, max_iter, extra_s=0.0):
    _, hist, ok = reconstruct(scattering, S, y0, tol=tol, max_iter=max_iter, method=method, lr=lr,
                              time_limit=time_limit)
    it, t, r = hist[-1]
    return {'method': method, 'lr': lr, 'reached': bool(ok), 'iterations': it, 'seconds': t + extra_s,
            'final_rel': r, 'time_to': time_to_levels(hist, extra_s),
            'curve': [(h[0], round(h[1], 4), h[2]) for h in hist[::max(1, len(hist) // 20)]]}


def compare(scattering, decoder, n=256, seed=99, tol=1e-3, method='lbfgs', lr=None, time_limit=30.0,
            max_iter=100000):
    """Time to tol from the warm start and from noise on n unseen chirps."""
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
    out = {'tol': tol, 'n_signals': n, 'loss_definition': LOSS_DEFINITION, 'method': method,
           'start_rel': {'plain': rel_error(scattering, noise, S), 'warm': rel_error(scattering, y0, S)},
           'decoder_forward_s': t_dec}
    out['warm'] = _run(scattering, S, y0, tol, method, lr, time_limit, max_iter, t_dec)
    out['plain'] = _run(scattering, S, noise, tol, method, lr, time_limit, max_iter)
    w, p = out['warm'], out['plain']
    out['speedup_by_level'] = {k: (p['time_to'][k]['seconds'] / w['time_to'][k]['seconds']
                                   if w['time_to'][k] and p['time_to'][k] else None) for k in w['time_to']}
    if w['reached'] and p['reached']:
        out['speedup'] = p['seconds'] / w['seconds']
    elif w['reached']:
        out['speedup_at_least'] = p['seconds'] / w['seconds']  # plain stopped at the time limit
        out['speedup'] = None
    else:
        out['speedup'] = None
    return out


def main(argv=None):
    from ..torch import Scattering1D
    ap = argparse.ArgumentParser()
    ap.add_argument('--checkpoint', default=CHECKPOINT)
    ap.add_argument('--tol', type=float, default=1e-3)
    ap.add_argument('--method', default='lbfgs')
    ap.add_argument('--lr', type=float, default=None)
    ap.add_argument('--time-limit', type=float, default=30.0)
    args = ap.parse_args(argv)
    dev = torch.device('cuda')
    sc = Scattering1D(6, 4096, 8).to(dev)
    dec = load_decoder(args.checkpoint, dev)
    print(json.dumps(compare(sc, dec, tol=args.tol, method=args.method, lr=args.lr, time_limit=args.time_limit)))


if __name__ == '__main__':
    main()
