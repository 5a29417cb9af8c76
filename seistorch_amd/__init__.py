"""seistorch_amd: MI355X-native finite-difference wave propagation.

A drop-in equation backend for seistorch's 2D hot path (acoustic and elastic
time stepping, PML/HABC damping, source injection, receiver sampling and the
adjoint gradient) built on hand-written HIP kernels for gfx950.  The kernels
live in ``libseistorch_amd.so`` (C-ABI, include/seistorch_amd.h); Python
mirrors the reference's plugin interface around them.
"""
__version__ = "0.1.0"
