"""Reconstruction of signals from scattering coefficients by descent through
the hand-written Scattering1D backward.

The precedent is examples/1d/reconstruct_torch.py:77-129: from y = randn,
descend on err = ||S(x) - S(y)|| (torch.norm, :104) with a bold-driver step
(:86-87, 117-125).  Here a batch of B signals is reconstructed at once and
the stopping criterion is the relative scattering error of the batch,

    rel(y) = ||S(y) - S_target||_F / ||S_target||_F        (not squared),

reached when rel(y) <= tol.  The signals are independent problems: each
descends on its own f_i(y_i) = ||S(y_i) - S_target,i||^2, whose gradients
one forward_train + backward of the batch gives together.  Methods:

  'lbfgs' (default)  limited-memory BFGS per signal (two-loop recursion on
                     each signal's own history), unit trial step with a
                     per-signal Armijo test; a signal whose step fails keeps
                     its point and halves its step.
  'adam'             Adam on y (kym_dec_adam), fixed step size.
  'bold'             gradient descent with the reference's bold driver, per
                     signal: step x1.1 after a decrease, x0.55 and the step
                     undone after an increase.

The start y0 is N(0, 1) noise for plain descent or the decoder's output for
the learned warm start.
"""
import ctypes
import time

import torch

from .. import _lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class ScatteringObjective:
    """f_i(y_i) = ||S(y_i) - S_target,i||^2 per signal and its gradient,
    from one kym_scat1d_forward_train + kym_scat1d_backward of the batch.
    Buffers are allocated once per batch size."""

    def __init__(self, scattering, S_target):
        self.sc = scattering
        self.dev = S_target.device
        self.dtype = S_target.dtype
        self.B = S_target.shape[0]
        self.target = S_target.reshape(self.B, -1).contiguous()
        self.plan = scattering._plan(self.dev, self.dtype)
        self.filt = scattering._packed_filters(self.plan, self.dev, self.dtype)
        self.out = torch.empty_like(self.target)
        self.ws = torch.empty(self.B * self.plan.train_ws_bytes_per_signal, dtype=torch.uint8, device=self.dev)
        self.target_sq = float((self.target.double() ** 2).sum().item())
        self.evaluations = 0

    def __call__(self, y, grad=True):
        """y (B, N) -> (f (B,), g (B, N) or None)."""
        y = y.contiguous()
        self.evaluations += 1
        with torch.cuda.device(self.dev):
            if grad:
                self.plan.forward_train(y, self.filt, self.out, self.B, self.ws)
            else:
                self.plan.forward(y, self.filt, self.out, self.B, self.ws)
            d = self.out - self.target
            f = (d * d).sum(dim=1)
            if not grad:
                return f, None
            g = torch.empty_like(y)
            d.mul_(2.0)
            self.plan.backward(d, self.filt, g, self.B, self.ws)
        return f, g

    def rel(self, f):
        """Relative scattering error of the batch from the per-signal f."""
        return (f.double().sum().item() / self.target_sq) ** 0.5


def rel_error(scattering, y, S_target):
    """||S(y) - S_target|| / ||S_target|| over the batch (the un-squared ratio)."""
    with torch.no_grad():
        return ((scattering(y) - S_target).double().norm() / S_target.double().norm()).item()


def _bdot(a, b):
    return (a * b).sum(dim=1)


def _lbfgs_direction(g, hist, gamma):
    """Two-loop recursion per signal; hist = [(s, y, rho)], rho = 0 marks a
    pair a signal does not use."""
    q = g.clone()
    alphas = []
    for s, yv, rho in reversed(hist):
        a = rho * _bdot(s, q)
        q.addcmul_(yv, a[:, None], value=-1.0)
        alphas.append(a)
    r = q * gamma[:, None]
    for (s, yv, rho), a in zip(hist, reversed(alphas)):
        b = rho * _bdot(yv, r)
        r.addcmul_(s, (a - b)[:, None])
    return r.neg_()


def reconstruct(scattering, S_target, y0, tol=1e-3, max_iter=1000, method='lbfgs', lr=None, time_limit=None,
                history=10, check_every=1, b1=0.9, b2=0.999, eps=1e-8):
    """Descend from y0 until rel(y) <= tol (or max_iter / time_limit seconds).

    Returns (y, hist, reached): hist holds (iteration, seconds, rel) at the
    checks; seconds count from the first evaluation, device work included."""
    obj = ScatteringObjective(scattering, S_target)
    dev = obj.dev
    y = y0.detach().reshape(obj.B, -1).clone().contiguous()
    hist = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    f, g = obj(y)
    reached = False

    def check(it):
        nonlocal reached
        r = obj.rel(f)
        hist.append((it, time.perf_counter() - t0, r))
        if r <= tol:
            reached = True
        return reached or (time_limit is not None and time.perf_counter() - t0 > time_limit)

    if method == 'adam':
        lr = 0.1 if lr is None else lr
        m, v = torch.zeros_like(y), torch.zeros_like(y)
        lib, s = _lib.lib(), _lib.stream_of(dev)
        for it in range(max_iter):
            if (it % check_every == 0) and check(it):
                break
            _lib.check(lib.kym_dec_adam(_p(y), _p(g.contiguous()), _p(m), _p(v), y.numel(), lr, b1, b2, eps, it + 1, s))
            f, g = obj(y)
        else:
            check(max_iter)
    elif method == 'bold':
        # examples/1d/reconstruct_torch.py:86-87: lr 100 on the un-squared
        # norm; on f = err^2 the gradient is 2 err times that one.
        step = torch.full((obj.B,), 100.0 if lr is None else lr, device=dev, dtype=y.dtype)
        for it in range(max_iter):
            if (it % check_every == 0) and check(it):
                break
            err = f.clamp_min(1e-30).sqrt()
            trial = y - (step / (2 * err))[:, None] * g
            f_t, g_t = obj(trial)
            better = f_t <= f
            step = torch.where(better, step * 1.1, step * 0.55)
            y = torch.where(better[:, None], trial, y)
            f = torch.where(better, f_t, f)
            g = torch.where(better[:, None], g_t, g)
        else:
            check(max_iter)
    elif method == 'lbfgs':
        pairs = []
        t = torch.ones(obj.B, device=dev, dtype=y.dtype)
        gamma = (0.1 * y.norm(dim=1) / g.norm(dim=1).clamp_min(1e-30))
        for it in range(max_iter):
            if (it % check_every == 0) and check(it):
                break
            d = _lbfgs_direction(g, pairs, gamma)
            gd = _bdot(g, d)
            # a direction that is not a descent direction restarts on -g
            bad = gd >= 0
            if bool(bad.any()):
                d = torch.where(bad[:, None], -g * gamma[:, None], d)
                gd = _bdot(g, d)
            trial = y + t[:, None] * d
            f_t, g_t = obj(trial)
            ok = f_t <= f + 1e-4 * t * gd
            s_k = trial - y
            y_k = g_t - g
            sy = _bdot(s_k, y_k)
            use = ok & (sy > 1e-10 * _bdot(s_k, s_k).sqrt() * _bdot(y_k, y_k).sqrt())
            rho = torch.where(use, 1.0 / torch.where(use, sy, torch.ones_like(sy)), torch.zeros_like(sy))
            okf = ok[:, None]
            pairs.append((torch.where(okf, s_k, torch.zeros_like(s_k)), torch.where(okf, y_k, torch.zeros_like(y_k)),
                          rho))
            if len(pairs) > history:
                pairs.pop(0)
            gamma = torch.where(use, sy / _bdot(y_k, y_k).clamp_min(1e-30), gamma)
            y = torch.where(okf, trial, y)
            f = torch.where(ok, f_t, f)
            g = torch.where(okf, g_t, g)
            t = torch.where(ok, torch.clamp(t * 2.0, max=1.0), t * 0.5)
        else:
            check(max_iter)
    else:
        raise ValueError("method must be 'lbfgs', 'adam' or 'bold'")
    torch.cuda.synchronize(dev)
    return y.reshape(y0.shape), hist, reached
