"""CPU oracle for seistorch's 2D finite-difference hot path (TEST INFRASTRUCTURE).

This package is the checker, never the product. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it. The product path (``seistorch_amd``) never imports anything from
here and fails loudly when its HIP library is missing.

Modules
-------
``acoustic``   FP64 numpy restatement of the constant-density acoustic PML
               step (seistorch/equations2d/acoustic.py:60-86), the time loop
               (seistorch/rnn.py:178-212), source injection
               (seistorch/source.py:47-57), receiver sampling
               (seistorch/probe.py:42-44), the L2 misfit
               (seistorch/loss.py:409-421) and the hand-derived exact discrete
               adjoint gradient of that loop.
``coeffs``     FP64 restatements of the PML profile (seistorch/pml.py:29-59),
               the HABC weights/masks (seistorch/habc.py:4-81) and the Ricker
               wavelet (seistorch/utils.py:235-249).
``elastic``    FP64 numpy restatement of the velocity-stress step
               (seistorch/equations2d/elastic.py:7-37,
               seistorch/equations2d/utils.py:3-48) and its exact adjoint.
``habc``       FP64 numpy restatement of the acoustic HABC step
               (seistorch/equations2d/acoustic_habc.py:79-221).
``ref_torch``  FP32 op-for-op torch restatement of the reference path (same
               conv2d Laplacian, same elementwise order, same masked source
               add and gather); it stands in for the reference PyTorch path
               and is the timed CPU baseline.

How each function is pinned (beyond being read off the reference) is listed
in the docstring of that function and exercised in ``tests/test_oracle_*.py``.
"""
