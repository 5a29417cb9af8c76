"""Host logic of the two-step adjoint's receiver map (CPU tensors)."""
import torch

from seistorch_amd.equations2d.acoustic import shared_receiver_map


def _flat(cells_per_shot, nz, nx):
    return torch.cat([b * nz * nx + torch.as_tensor(c, dtype=torch.int64)
                      for b, c in enumerate(cells_per_shot)])


def test_shared_cells_give_the_index_map():
    nz, nx = 6, 9
    cells = [2 * nx + 1, 2 * nx + 4, 3 * nx + 7]
    ok, m = shared_receiver_map(_flat([cells] * 3, nz, nx), 3, nz, nx)
    assert ok
    assert m.shape == (nz, nx) and m.dtype == torch.int32
    flat = m.view(-1)
    assert [int(flat[c]) for c in cells] == [0, 1, 2]
    assert int((flat >= 0).sum()) == 3


def test_no_receivers():
    ok, m = shared_receiver_map(torch.empty(0, dtype=torch.int64), 2, 5, 5)
    assert ok and m is None


def test_rejects_differing_cells_duplicates_and_ragged_counts():
    nz, nx = 6, 9
    a, b = [10, 11, 12], [10, 11, 13]
    assert not shared_receiver_map(_flat([a, b], nz, nx), 2, nz, nx)[0]
    assert not shared_receiver_map(_flat([[10, 10, 12]] * 2, nz, nx), 2, nz, nx)[0]
    assert not shared_receiver_map(_flat([[10, 11], [10, 11, 12]], nz, nx), 2, nz, nx)[0]
    # same cells in another order per shot
    assert not shared_receiver_map(_flat([[10, 11], [11, 10]], nz, nx), 2, nz, nx)[0]
