"""kym_scat2d_forward_noalloc (the entry point that allocates nothing; its
workspace is a static device array of KYM_NOALLOC_SCRATCH_BYTES) against
kym_scat2d_forward given a workspace of the same size: bit-identical outputs
on two calls in a row, the second with a smaller batch, and a clear error
for a plan whose one image needs more than the static array."""
import ctypes

import numpy as np
import pytest
import torch

from kymatio_amd import _lib
from kymatio_amd.scattering2d.frontend.plan import Scat2DPlanHandle
from kymatio_amd.torch import Scattering2D

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda')
SCRATCH = 64 << 20  # KYM_NOALLOC_SCRATCH_BYTES


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _noalloc(plan, x, filt, out, batch):
    lib = _lib.lib()
    return lib.kym_scat2d_forward_noalloc(plan._h, _vp(x), _vp(filt), _vp(out), int(batch), _lib.stream_of(x.device))


def test_noalloc_matches_forward_twice_then_smaller():
    sc = Scattering2D(3, (256, 256)).to(DEV)
    plan = sc._plan(DEV, torch.float32)
    filt = sc._packed_filters(plan, DEV, torch.float32)
    assert 0 < plan.ws_bytes_per_image <= SCRATCH
    per_chunk = SCRATCH // plan.ws_bytes_per_image
    g = torch.Generator(device=DEV).manual_seed(3)
    ws = torch.empty(SCRATCH, dtype=torch.uint8, device=DEV)
    for batch in (2 * per_chunk + 3, 5):  # several chunks, then fewer images than one chunk
        x = torch.randn(batch, 256, 256, device=DEV, generator=g)
        ref = torch.full((batch, plan.C, plan.oh, plan.ow), float('nan'), device=DEV)
        got = torch.full_like(ref, float('nan'))
        plan.forward(x, filt, ref, batch, ws)
        _lib.check(_noalloc(plan, x, filt, got, batch))
        torch.cuda.synchronize()
        print('batch %d (%d images per chunk): max |noalloc - forward| = %.3g'
              % (batch, per_chunk, (got - ref).abs().max().item()))
        assert torch.isfinite(got).all()
        assert torch.equal(got, ref)


def test_noalloc_rejects_plans_beyond_the_scratch():
    big = None
    for side, J in ((512, 6), (512, 5)):  # 640x640 and 576x576 padded, deep trees
        p = Scat2DPlanHandle(side, side, J, 8, 2, 0, torch.float32)
        print('%dx%d J=%d: %d workspace bytes per image' % (side, side, J, p.ws_bytes_per_image))
        if p.ws_bytes_per_image > SCRATCH:
            big = p
            break
    assert big is not None, 'no test size needs more than the scratch'
    x = torch.zeros(1, 4, device=DEV)  # never read: the size check comes first
    rc = _noalloc(big, x, x, x, 1)
    assert rc == 3
    msg = _lib.lib().kym_last_error().decode()
    assert 'static scratch' in msg and str(SCRATCH) in msg, msg
    with pytest.raises(ValueError, match='static scratch'):
        _lib.check(rc)
