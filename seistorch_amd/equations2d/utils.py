"""Ring save/restore of the boundary-saving scheme (host side).

The ring is the band of N cells just inside the damped strips of a
(B, nz, nx) field: rows [P, P+N) and [nz-P-N, nz-P) across the whole width,
columns [P, P+N) and [nx-P-N, nx-P) down the whole height, P = NPML.  It is
the interface of seistorch/equations2d/utils.py:50-89: the tuple
(top, bottom, left, right) that save_boundaries returns is what the
reference's checkpointing hands to ``_time_step_backward``.  Restoring
writes the four bands in that order (later bands win at the four crossing
cells, which hold equal values when the tuple comes from one field); with a
free surface (``multiple``) the top band is not restored.  P = 0 gives the
outermost band.  The fused HIP loops keep the same ring inside the library;
these helpers serve the per-step interface.
"""
import torch


def _bands(nz, nx, NPML, N):
    """Index ranges (start, stop) of the top, bottom, left and right bands."""
    return ((NPML, NPML + N), (nz - NPML - N, nz - NPML),
            (NPML, NPML + N), (nx - NPML - N, nx - NPML))


def save_boundaries(tensor: torch.Tensor, NPML: int = 49, N: int = 1):
    """(top, bottom, left, right): copies of the ring bands of `tensor`."""
    nz, nx = tensor.shape[-2:]
    (t0, t1), (b0, b1), (l0, l1), (r0, r1) = _bands(nz, nx, NPML, N)
    return (tensor[..., t0:t1, :].clone(), tensor[..., b0:b1, :].clone(),
            tensor[..., :, l0:l1].clone(), tensor[..., :, r0:r1].clone())


def restore_boundaries(tensor, memory, NPML=49, N=1, multiple=False):
    """Write the saved bands back into `tensor` (in place) and return it."""
    top, bottom, left, right = memory
    nz, nx = tensor.shape[-2:]
    (t0, t1), (b0, b1), (l0, l1), (r0, r1) = _bands(nz, nx, NPML, N)
    if not multiple:
        tensor[..., t0:t1, :] = top
    tensor[..., b0:b1, :] = bottom
    tensor[..., :, l0:l1] = left
    tensor[..., :, r0:r1] = right
    return tensor
