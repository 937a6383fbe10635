"""CPU oracle for the scattering hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, the algorithm of kymatio's torch backend
for Scattering2D / Scattering1D (filter banks, the scattering cores and the
backend ops), so that the HIP implementation in ``kymatio_amd`` can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or
as the timed CPU baseline.  The product (``kymatio_amd``) never imports it.

Modules
-------
filters2d   literal restatement of kymatio/scattering2d/filter_bank.py
filters1d   restatement of kymatio/scattering1d/filter_bank.py + frontend setup
scat2d      op-for-op restatement of the torch backend 2D path
            (torch CPU, FP32 "as reference") and of the NumPy backend (FP64)
scat1d      the same for the 1D path

How each function is pinned (checked against something other than itself):
* filters2d, scat2d  -- reference fixture tests/scattering2d/test_data_2d.npz
  (J=4, 32x32, P=64) end to end; closed-form checks of padding, periodization
  and modulus; the golden vectors written by tests/golden/make_golden.py from
  the reference at P=40/160/272.
* filters1d, scat1d  -- reference fixture tests/scattering1d/test_data_1d.npz
  (J=6, Q=16, N=512); golden vectors at J=6, Q=8, N=4096.
"""
