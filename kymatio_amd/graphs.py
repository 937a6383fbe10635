"""Replay of a whole transform as one captured HIP graph.

Transforms that run as many small HIP launches (time-frequency scattering
calls a few thousand kernels per batch) are bound by the host: each launch
costs far more on the CPU than on the GPU.  GraphCache captures the
function once per input signature into a HIP graph (torch.cuda.graph, whose
capture records the library's kernels launched on the capturing stream)
and replays it: one host call per batch.  The input is copied into the
graph's static input, the outputs are copied out of its static outputs.
"""
import torch


def _clone(out):
    if torch.is_tensor(out):
        return out.clone()
    if isinstance(out, dict):
        return {k: _clone(v) for k, v in out.items()}
    if isinstance(out, (list, tuple)):
        return type(out)(_clone(v) for v in out)
    return out


class GraphCache:
    """fn(x) -> tensor / dict / list of tensors, captured per key."""

    def __init__(self, max_entries=8):
        self._entries = {}
        self.max_entries = max_entries
        self.captures = 0

    def clear(self):
        self._entries.clear()

    def __call__(self, key, fn, x):
        entry = self._entries.get(key)
        if entry is None:
            if len(self._entries) >= self.max_entries:
                self._entries.pop(next(iter(self._entries)))
            dev = x.device
            static_x = x.detach().clone()
            cur = torch.cuda.current_stream(dev)
            side = torch.cuda.Stream(dev)
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                fn(static_x)  # warm-up outside the capture (plans, filters, allocator)
            cur.wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                out = fn(static_x)
            entry = (g, static_x, out)
            self._entries[key] = entry
            self.captures += 1
        g, static_x, out = entry
        static_x.copy_(x)
        g.replay()
        return _clone(out)
