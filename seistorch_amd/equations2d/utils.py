"""Ring save/restore of the boundary-saving scheme (host side).

Mirrors seistorch/equations2d/utils.py:50-90: the 1-cell ring at index NPML
(rows NPML and -(NPML+1), columns NPML and -(NPML+1)) of a (B, nz, nx) field.
The fused HIP loop stores the same ring inside the library; these helpers
serve the per-step ``_time_step_backward`` interface.
"""
import torch


def save_boundaries(tensor: torch.Tensor, NPML: int = 49, N: int = 1):
    top = tensor[:, NPML:NPML + N, :].clone()
    bottom = tensor[:, -N:, :].clone() if NPML == 0 else tensor[:, -(NPML + N):-NPML, :].clone()
    left = tensor[..., NPML:NPML + N].clone()
    right = tensor[..., -N:].clone() if NPML == 0 else tensor[..., -(NPML + N):-NPML].clone()
    return top, bottom, left, right


def restore_boundaries(tensor, memory, NPML=49, N=1, multiple=False):
    top, bottom, left, right = memory
    if not multiple:
        tensor[:, NPML:NPML + N, :] = top
    if NPML != 0:
        tensor[:, -(NPML + N):-NPML, :] = bottom
    else:
        tensor[:, -N:, :] = bottom
    tensor[..., NPML:NPML + N] = left
    if NPML != 0:
        tensor[..., -(NPML + N):-NPML] = right
    else:
        tensor[..., :, -N:] = right
    return tensor
