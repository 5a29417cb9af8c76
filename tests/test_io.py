"""On-disk formats (seistorch_amd/io.py): geometry pickles, shot records,
models and the coordinate padding, on the CPU."""
import os
import pickle

import numpy as np
import pytest

from seistorch_amd import io as sio
from seistorch_amd import synthetic


def _geometry(nshots=3, nx=40):
    sources = [[5 + 10 * i, 1] for i in range(nshots)]
    receivers = [[list(range(0, nx, 2)), [3] * (nx // 2)] for _ in range(nshots)]
    receivers[1] = [list(range(1, nx, 3)), [4] * len(range(1, nx, 3))]  # a towed shot
    return sources, receivers


def test_geometry_pickles_round_trip(tmp_path):
    sources, receivers = _geometry()
    sio.write_pkl(tmp_path / "sources.pkl", sources)
    sio.write_pkl(tmp_path / "receivers.pkl", receivers)
    assert sio.read_pkl(tmp_path / "sources.pkl") == sources
    assert sio.read_pkl(tmp_path / "receivers.pkl") == receivers
    # numpy scalars and arrays inside the lists (as np.arange(...).tolist() is not always used)
    mixed = [[np.int64(3), np.float32(2.0)], [np.arange(4), np.zeros(4)]]
    sio.write_pkl(tmp_path / "m.pkl", mixed)
    back = sio.read_pkl(tmp_path / "m.pkl")
    assert back[0] == [3, 2.0] and np.array_equal(back[1][0], np.arange(4))


class _Planted:
    """A pickle whose loading would call a function: the restricted loader
    must refuse it before anything runs."""

    def __reduce__(self):
        return (os.getcwd, ())


def test_pickled_code_is_refused(tmp_path):
    p = tmp_path / "evil.pkl"
    with open(p, "wb") as f:
        pickle.dump([1, 2, _Planted()], f)
    with pytest.raises(pickle.UnpicklingError, match="refusing"):
        sio.read_pkl(p)
    arr = np.empty(2, dtype=object)
    arr[0], arr[1] = np.zeros(3), _Planted()
    np.save(tmp_path / "evil.npy", arr, allow_pickle=True)
    with pytest.raises(pickle.UnpicklingError, match="refusing"):
        sio.load_npy(tmp_path / "evil.npy")


def test_records_npy_object_and_regular(tmp_path):
    rng = np.random.default_rng(0)
    shots = [rng.standard_normal((50, nr, 1)).astype(np.float32) for nr in (20, 13, 20)]
    rec = np.empty(3, dtype=np.ndarray)  # io.py:130-133
    for i, s in enumerate(shots):
        rec[i] = s
    np.save(tmp_path / "obs.npy", rec)
    dl = sio.DataLoader(str(tmp_path / "obs.npy"))
    assert dl.shape == (3,)
    for i in range(3):
        assert np.array_equal(dl[i], shots[i])
    reg = rng.standard_normal((2, 50, 7, 2)).astype(np.float32)
    np.save(tmp_path / "reg.npy", reg)
    assert np.array_equal(sio.DataLoader(str(tmp_path / "reg.npy"))[1], reg[1])
    with pytest.raises(RuntimeError, match="h5py"):
        sio.DataLoader(str(tmp_path / "obs.hdf5"))


def test_seisrecord_forward_then_inversion(tmp_path):
    sources, receivers = _geometry()
    sio.write_pkl(tmp_path / "receivers.pkl", receivers)
    cfg = {"geom": {"obsPath": str(tmp_path / "obs.npy"),
                    "receivers": str(tmp_path / "receivers.pkl")}}
    w = sio.SeisRecord(cfg)
    w.setup("forward")
    data = [np.full((30, len(r[0]), 1), i, np.float32) for i, r in enumerate(receivers)]
    for i, d in enumerate(data):
        w[i] = d
    w.write()
    r = sio.SeisRecord(cfg)
    r.setup("inversion")
    assert r.shape == (3,)
    for i in range(3):
        assert np.array_equal(r[i], data[i])


def test_coordinate_padding_known_answers():
    """setup.py:386-483: +width on both axes; with multiples no top padding
    (the source's check is on the padded z, so z = 0 is still unpadded)."""
    assert sio.setup_src_coords([10, 3], 50) == (53, 60)
    assert sio.setup_src_coords([10, 3], 50, multiple=True) == (3, 60)
    assert sio.setup_src_coords([10, 0], 50, multiple=True) == (0, 60)
    assert sio.setup_src_coords([10.7, 3.2], 50) == (53, 60)
    r = sio.setup_rec_coords([[0, 2, 4], [5, 5, 6]], 50)
    assert r.tolist() == [[55, 50], [55, 52], [56, 54]]
    r = sio.setup_rec_coords([[0, 2], [5, 5]], 50, multiple=True)
    assert r.tolist() == [[5, 50], [5, 52]]
    with pytest.raises(TypeError):
        sio.setup_src_coords([[1], 2], 50)


def test_acquisition_feeds_the_propagation_layout(tmp_path):
    """Files -> padded sources and receivers -> flat receiver indices in
    shot order; split_records inverts the shot stacking (rnn.py:207-212)."""
    sources, receivers = _geometry()
    sio.write_pkl(tmp_path / "s.pkl", sources)
    sio.write_pkl(tmp_path / "r.pkl", receivers)
    cfg = {"geom": {"sources": str(tmp_path / "s.pkl"), "receivers": str(tmp_path / "r.pkl"),
                    "boundary": {"width": 50}, "multiple": False}}
    src, rec = sio.acquisition_from_cfg(cfg)
    assert src.dtype == np.int32 and src.tolist() == [[51, 55], [51, 65], [51, 75]]
    nz, nx = 130, 140
    flat = synthetic.rec_flat_index(rec, nz, nx)
    counts = [len(r) for r in rec]
    assert len(flat) == sum(counts)
    b1 = flat[counts[0]:counts[0] + counts[1]]
    assert np.array_equal(b1 // (nz * nx), np.ones(counts[1]))
    assert np.array_equal((b1 % (nz * nx)) // nx, rec[1][:, 0])
    stacked = np.arange(10 * sum(counts), dtype=np.float32).reshape(10, -1)
    shots = sio.split_records(stacked, counts)
    assert [s.shape for s in shots] == [(10, c, 1) for c in counts]
    assert np.array_equal(sio.stack_records(shots)[..., 0], stacked)


def test_model_npy(tmp_path):
    m = np.linspace(1500, 3000, 12).reshape(3, 4)
    np.save(tmp_path / "vp.npy", m)
    v = sio.load_model(str(tmp_path / "vp.npy"))
    assert v.dtype == np.float32 and np.array_equal(v, m.astype(np.float32))
