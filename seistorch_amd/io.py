"""On-disk formats of seistorch's YAML-driven runs, read and written without
executing anything from the files.

Mirrors the parts of seistorch/io.py and seistorch/setup.py that turn the
files a configuration names into the inputs of the propagation path:
* geometry: ``sources.pkl`` = one ``[x, z]`` per shot, ``receivers.pkl`` =
  one ``[[x...], [z...]]`` per shot, in grid units of the unpadded model
  (doc/data_format.md:55-85; read with ``read_pkl``, io.py:371 and
  utils.py:195-199);
* coordinates: +boundary width on both axes, no top padding of z when
  multiples are modelled (setup.py:386-483, ``setup_src_coords`` /
  ``setup_rec_coords``);
* shot records: ``.npy`` holding an object array of per-shot
  ``(nt, nr, nc)`` float32 arrays, or a regular ``(nshots, nt, nr, nc)``
  array (io.py:27-60, 84-224: ``DataLoader``, ``SeisRecord``);
* models: ``.npy`` ``(nz, nx)`` arrays.
HDF5 records (``shot_{i}`` datasets) need h5py, which is not installed
here: those paths raise with that reason.

The reference unpickles with ``pickle.load`` / ``np.load(allow_pickle=True)``;
here pickled payloads go through an unpickler that admits only plain
containers and numpy's array reconstruction, so a file cannot run code.
"""
import io as _io
import os
import pickle

import numpy as np

# the globals a pickled list of numbers or a numpy object array of float
# arrays refers to; nothing else is resolved
_ALLOWED = {
    ("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"), ("numpy._core.multiarray", "scalar"),
    ("numpy", "ndarray"), ("numpy", "dtype"),
}


class _SafeUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _ALLOWED:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(f"refusing to load {module}.{name} from a data file")


def _safe_loads(data):
    return _SafeUnpickler(_io.BytesIO(data)).load()


def read_pkl(path):
    """utils.py:195-199 / io.py:371 (doc/data_format.md:47-51), restricted
    to data."""
    with open(path, "rb") as f:
        return _SafeUnpickler(f).load()


def write_pkl(path, data):
    """utils.py:336-339 / io.py:501-504."""
    with open(path, "wb") as f:
        pickle.dump(data, f)


def load_npy(path):
    """An .npy file: a plain array with np.load(allow_pickle=False), an
    object array (the records of shots with different receiver counts)
    through the restricted unpickler."""
    with open(path, "rb") as f:
        version = np.lib.format.read_magic(f)
        if version == (1, 0):
            shape, fortran, dtype = np.lib.format.read_array_header_1_0(f)
        else:
            shape, fortran, dtype = np.lib.format.read_array_header_2_0(f)
        if not dtype.hasobject:
            return np.load(path, allow_pickle=False)
        arr = _safe_loads(f.read())
    if not isinstance(arr, np.ndarray) or arr.shape != tuple(shape):
        raise ValueError(f"{path}: object payload does not match its header {shape}")
    return arr


def load_model(path):
    """A model file (nz, nx) as float32."""
    m = load_npy(path)
    if m.dtype == object or m.ndim != 2:
        raise ValueError(f"{path}: expected a 2-D numeric model, got {m.dtype} {m.shape}")
    return np.ascontiguousarray(m, np.float32)


# ------------------------------------------------------------ coordinates
def setup_src_coords(coords, bwidth, multiple=False):
    """setup.py:452-483 for 2-D: coords = [x, z] -> padded (z, x).  The
    reference asserts int or float coordinates and indexes with their
    integer part."""
    if len(coords) != 2:
        raise NotImplementedError("only 2-D sources [x, z] are supported")
    for v in coords:
        if not isinstance(v, (int, float, np.integer, np.floating)):
            raise TypeError(f"the source location must be a number, got {type(v)}")
    x, z = coords
    x, z = x + bwidth, z + bwidth
    if multiple and bool(z):  # setup.py:477: the check is on the padded value
        z -= bwidth
    return int(z), int(x)


def setup_rec_coords(coords, bwidth, multiple=False):
    """setup.py:386-415 for 2-D: coords = [[x...], [z...]] -> (nrec, 2)
    padded (z, x)."""
    if len(coords) != 2:
        raise NotImplementedError("only 2-D receivers [[x...], [z...]] are supported")
    xs = [v + bwidth for v in coords[0]]
    zs = [v + bwidth for v in coords[1]]
    if multiple:
        zs = [v - bwidth for v in zs]
    if len(xs) != len(zs):
        raise ValueError("receiver x and z lists differ in length")
    return np.stack([np.asarray(zs), np.asarray(xs)], 1).astype(np.int64)


def setup_acquisition(src_list, rec_list, bwidth, multiple=False, shots=None):
    """setup.py:261-289: the padded sources (B, 2) int32 (z, x) and the
    per-shot receiver arrays (nrec_b, 2) (z, x) of the chosen shots, ready
    for ``synthetic.rec_flat_index`` and ``equations2d.acoustic.propagate``."""
    if len(src_list) != len(rec_list):
        raise ValueError("the lengths of sources and receivers must be equal")
    shots = range(len(src_list)) if shots is None else shots
    src = np.array([setup_src_coords(src_list[i], bwidth, multiple) for i in shots], np.int32)
    rec = [setup_rec_coords(rec_list[i], bwidth, multiple) for i in shots]
    return src.reshape(-1, 2), rec


def acquisition_from_cfg(cfg, shots=None):
    """Sources and receivers named by a configuration (geom.sources,
    geom.receivers, geom.boundary.width, geom.multiple; setup.py:417-450)."""
    geom = cfg["geom"]
    src_list = read_pkl(geom["sources"])
    rec_list = read_pkl(geom["receivers"])
    return setup_acquisition(src_list, rec_list, geom["boundary"]["width"],
                             geom.get("multiple", False), shots)


# ---------------------------------------------------------------- records
def split_records(rec, counts):
    """rnn.py:207-212: the stacked records (nt, sum(counts)[, nch]) of
    propagate -> one (nt, nr_b, nch) array per shot."""
    rec = np.asarray(rec)
    if rec.ndim == 2:
        rec = rec[..., None]
    if rec.shape[1] != int(np.sum(counts)):
        raise ValueError(f"records have {rec.shape[1]} traces, the receivers {np.sum(counts)}")
    edges = np.cumsum([0] + list(counts))
    return [np.ascontiguousarray(rec[:, edges[i]:edges[i + 1]], np.float32)
            for i in range(len(counts))]


def stack_records(shots):
    """The inverse of split_records: (nt, sum nr_b, nch)."""
    return np.concatenate([np.asarray(s, np.float32) for s in shots], 1)


def _ext(path):
    return os.path.splitext(path)[1]


def _no_hdf5(path):
    raise RuntimeError(f"{path}: HDF5 records need h5py, which is not installed; "
                       "use .npy records")


class DataLoader:
    """io.py:27-84 (.npy only)."""

    def __init__(self, dpath, dkey="shot"):
        self.dpath = dpath
        self.dkey = dkey
        self.filetype = _ext(dpath)
        if self.filetype in (".hdf5", ".h5"):
            _no_hdf5(dpath)
        if self.filetype != ".npy":
            raise ValueError(f"{dpath}: unsupported record format {self.filetype!r}")
        self.data = load_npy(dpath)

    def __getitem__(self, key):
        if isinstance(key, list):
            return [self.data[k] for k in key]
        return self.data[key]

    @property
    def shape(self):
        return self.data.shape


class SeisRecord:
    """io.py:86-224 (.npy only): the observed or synthetic shot records of a
    run.  cfg: geom.obsPath, geom.receivers."""

    def __init__(self, cfg, logger=None):
        self.cfg = cfg
        self.logger = logger
        self.datapath = cfg["geom"]["obsPath"]
        self.recpath = cfg["geom"]["receivers"]
        self.nshots = len(read_pkl(self.recpath))
        self.filetype = _ext(self.datapath)
        self.record = None

    def load_receivers(self, path=None):
        return read_pkl(self.recpath if path is None else path)

    def setup(self, mode):
        if self.filetype in (".hdf5", ".h5"):
            _no_hdf5(self.datapath)
        if mode == "forward":
            self.record = np.empty(self.nshots, dtype=np.ndarray)
        elif mode == "inversion":
            self.record = load_npy(self.datapath)
        else:
            raise ValueError(f"unknown mode {mode!r}")

    def write_shot(self, shot_no, data, datapath=None):
        self.record[shot_no] = data

    def write(self):
        np.save(self.datapath, self.record)

    @property
    def shape(self):
        return (self.record.shape[0],)

    def __setitem__(self, key, value):
        if isinstance(key, (int, np.integer)) and isinstance(value, np.ndarray):
            self.write_shot(int(key), value)
        else:
            raise NotImplementedError("The key must be an integer and the value must be a "
                                      "numpy array.")

    def __getitem__(self, key):
        if isinstance(key, list):
            return [self.record[k] for k in key]
        return self.record[key]
