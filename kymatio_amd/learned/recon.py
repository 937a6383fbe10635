"""Reconstruction of signals from scattering coefficients by descent through
the hand-written scattering backward (the precedent is
examples/1d/reconstruct_torch.py:77-129, which descends on ||S(y) - S(x)||).

The loss is the relative scattering loss over the batch,
    L(y) = ||S(y) - S_target||_F^2 / ||S_target||_F^2
(the squared relative error), minimised with Adam (kym_dec_adam) from an
initial signal: N(0, 1) noise for plain descent, or the decoder's output for
the learned warm start.
"""
import ctypes
import time

import torch

from .. import _lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def rel_loss(scattering, y, S_target):
    """L(y) = ||S(y) - S_target||^2 / ||S_target||^2."""
    with torch.no_grad():
        return (((scattering(y) - S_target).norm() / S_target.norm()) ** 2).item()


def reconstruct(scattering, S_target, y0, tol=1e-3, max_iter=1000, lr=1e-2, time_limit=None, b1=0.9, b2=0.999,
                eps=1e-8, check_every=1):
    """Adam on y from y0 until L(y) <= tol (or max_iter / time_limit seconds).

    Returns (y, history of (iteration, seconds, L), reached)."""
    lib = _lib.lib()
    dev = y0.device
    s = _lib.stream_of(dev)
    y = y0.detach().clone().contiguous()
    m = torch.zeros_like(y)
    v = torch.zeros_like(y)
    denom = S_target.norm() ** 2
    hist = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    reached = False
    for it in range(1, max_iter + 1):
        yy = y.detach().requires_grad_()
        diff = scattering(yy) - S_target
        loss = (diff * diff).sum() / denom
        loss.backward()
        if it == 1 or it % check_every == 0:
            r = loss.detach().item()
            hist.append((it - 1, time.perf_counter() - t0, r))
            if r <= tol:
                reached = True
                break
            if time_limit is not None and time.perf_counter() - t0 > time_limit:
                break
        g = yy.grad.contiguous()
        _lib.check(lib.kym_dec_adam(_p(y), _p(g), _p(m), _p(v), y.numel(), lr, b1, b2, eps, it, s))
    torch.cuda.synchronize(dev)
    return y, hist, reached
