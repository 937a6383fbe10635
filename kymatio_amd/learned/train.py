"""Training of the warm-start decoder on pairs (x, S(x)) generated on the fly.

Signals are chirps + noise made on the device (kym_chirps); S is the fused
Scattering1D.  The decoder is trained to make the scattering of its output
match: the loss is the relative scattering error of the batch,
    L = ||S(dec(S(x))) - S(x)||_F^2 / ||S(x)||_F^2,
back-propagated through the hand-written Scattering1D backward (the chirps'
phases are random, so a signal-space loss would only learn their mean).
"""
import ctypes
import time

import torch

from .. import _lib
from .decoder import ScatteringDecoder1D


def chirps(n, length, seed, index0, noise=0.05, device='cuda'):
    x = torch.empty(n, length, device=device)
    _lib.check(_lib.lib().kym_chirps(ctypes.c_void_p(x.data_ptr()), n, length, int(seed), int(index0), float(noise),
                                     _lib.stream_of(x.device)))
    return x


def scattering_loss_grad(scattering, y, S_target):
    """Relative scattering loss and its gradient w.r.t. y (through the HIP backward)."""
    yy = y.detach().requires_grad_()
    d = scattering(yy) - S_target
    loss = (d * d).sum() / (S_target * S_target).sum()
    loss.backward()
    return loss.detach(), yy.grad


def train(scattering, n_pairs=1_000_000, batch=256, lr=1e-3, seed=1234, log_every=200, time_limit=None,
          decoder=None, lr_final=1e-4, log=print):
    """Train (or continue training) a decoder on n_pairs generated pairs;
    returns (decoder, history of (step, pairs, seconds, batch loss))."""
    dev = next(scattering.buffers()).device
    N = scattering.shape[0]
    x0 = chirps(batch, N, seed, 0, device=dev)
    with torch.no_grad():
        S0 = scattering(x0)
    C, T = S0.shape[-2], S0.shape[-1]
    dec = decoder or ScatteringDecoder1D(C, T, N, device=dev, seed=seed)
    if decoder is None:
        dec.fit_scale(S0)
    steps = max(1, n_pairs // batch)
    hist = []
    t0 = time.perf_counter()
    for step in range(steps):
        # cosine decay from lr to lr_final
        frac = step / max(steps - 1, 1)
        lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + __import__('math').cos(3.141592653589793 * frac))
        x = chirps(batch, N, seed, (step + 1) * batch, device=dev)
        with torch.no_grad():
            S = scattering(x)
        y, acts = dec.forward(S, keep=True)
        loss, dy = scattering_loss_grad(scattering, y, S)
        grads = dec.backward(dy.contiguous(), acts)
        dec.adam_step(grads, lr=lr_t)
        if step % log_every == 0 or step == steps - 1:
            l = loss.item()
            el = time.perf_counter() - t0
            hist.append((step, (step + 1) * batch, el, l))
            log('step %6d  pairs %8d  %7.1f s  relative scattering loss %.4e' % (step, (step + 1) * batch, el, l))
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            break
    return dec, hist


CHECKPOINT = __import__('os').path.join(__import__('os').path.dirname(__import__('os').path.abspath(__file__)),
                                        'decoder_scat1d_j6q8_n4096.pt')


def main(argv=None):
    """python -m kymatio_amd.learned.train [--pairs N] [--lr LR] [--out PATH]:
    trains the warm-start decoder for Scattering1D J=6 Q=8 N=4096 on the GPU."""
    import argparse
    from ..torch import Scattering1D
    ap = argparse.ArgumentParser()
    ap.add_argument('--pairs', type=int, default=1_000_000)
    ap.add_argument('--batch', type=int, default=256)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--lr-final', type=float, default=1e-4)
    ap.add_argument('--seed', type=int, default=1234)
    ap.add_argument('--out', default=CHECKPOINT)
    args = ap.parse_args(argv)
    sc = Scattering1D(6, 4096, 8).to(torch.device('cuda'))
    dec, hist = train(sc, n_pairs=args.pairs, batch=args.batch, lr=args.lr, lr_final=args.lr_final, seed=args.seed)
    state = dec.state_dict()
    state['train_pairs'] = torch.tensor(args.pairs)
    state['train_final_loss'] = torch.tensor(hist[-1][3])
    torch.save(state, args.out)
    print('saved', args.out)


if __name__ == '__main__':
    main()
