"""Oracle (test infrastructure only): the learned decoder of
kymatio_amd/learned/decoder.py restated with torch CPU conv1d, so the MFMA
forward and hand-written backward kernels can be checked.

The GPU kernels feed bf16-rounded operands to the MFMA (weights, layer
inputs, upstream gradients) and accumulate in FP32; the oracle rounds the
weights and layer inputs the same way (round to nearest even, as
torch.bfloat16 does) with a straight-through gradient, and computes in FP64.
"""
import torch
import torch.nn.functional as F


def _bf16(t):
    return t + (t.to(torch.bfloat16).to(t.dtype) - t).detach()


def layer(x, L, scale=None):
    """One decoder layer on x (B, Cin, T_in): optional per-channel input
    scale (applied in FP32, as the kernel does, before the bf16 rounding),
    nearest upsampling by L['up'], 'same' convolution of kernel 9, bias,
    optional ReLU."""
    if scale is not None:
        x = (x.float() * scale.float()[None, :, None]).double() if not x.requires_grad else \
            x.double() * scale.double()[None, :, None]
    x = x.double()
    if L['up'] > 1:
        x = x.repeat_interleave(L['up'], dim=-1)
    w = _bf16(L['w'].double()).permute(0, 2, 1)  # (Cout, Cin, 9)
    x = F.conv1d(_bf16(x), w, L['b'].double(), padding=4)
    if L['relu']:
        x = torch.relu(x)
    return x


def forward(S, layers, scale):
    """layers: list of dicts with w (Cout, 9, Cin), b (Cout), up, relu.
    Returns the output and the per-layer outputs."""
    x = S
    outs = []
    for i, L in enumerate(layers):
        x = layer(x, L, scale if i == 0 else None)
        outs.append(x)
    return x[:, 0, :], outs
