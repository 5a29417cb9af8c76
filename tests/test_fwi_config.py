"""The configuration-driven FWI set-up (fwi.problem_from_config) on the CPU:
files named by a reference-style configuration -> the Problem's tensors."""
import numpy as np
import pytest

from seistorch_amd import fwi
from seistorch_amd import io as sio
from seistorch_amd.utils import ricker_wave


def _write_case(tmp_path, nz=30, nx=40, nt=50, nshots=2):
    rng = np.random.default_rng(0)
    vt = 2000.0 + 500.0 * rng.random((nz, nx))
    np.save(tmp_path / "true.npy", vt)
    np.save(tmp_path / "init.npy", np.full((nz, nx), 2200.0))
    sources = [[5 + 20 * i, 1] for i in range(nshots)]
    receivers = [[list(range(0, nx, 2)), [2] * (nx // 2)] for _ in range(nshots)]
    sio.write_pkl(tmp_path / "s.pkl", sources)
    sio.write_pkl(tmp_path / "r.pkl", receivers)
    obs = np.empty(nshots, dtype=np.ndarray)
    for i in range(nshots):
        obs[i] = rng.standard_normal((nt, nx // 2, 1)).astype(np.float32)
    np.save(tmp_path / "obs.npy", obs)
    cfg = {"equation": "acoustic",
           "geom": {"truePath": {"vp": str(tmp_path / "true.npy")},
                    "initPath": {"vp": str(tmp_path / "init.npy")},
                    "sources": str(tmp_path / "s.pkl"), "receivers": str(tmp_path / "r.pkl"),
                    "obsPath": str(tmp_path / "obs.npy"), "wavelet": None, "fm": 10.0,
                    "dt": 1e-3, "nt": nt, "h": 10.0, "wavelet_delay": 20,
                    "wavelet_inverse": False, "multiple": False,
                    "boundary": {"type": "pml", "width": 50}},
           "training": {"filter_ord": 3}}
    return cfg, obs


def test_problem_from_config_files(tmp_path):
    cfg, obs = _write_case(tmp_path)
    p = fwi.problem_from_config(cfg, device="cpu")
    assert tuple(p.vp_init.shape) == (130, 140) and float(p.vp_init[0, 0]) == 2200.0
    assert np.array_equal(p.wavelet.numpy(), ricker_wave(10.0, 1e-3, 50, 20).numpy())
    assert p.src.tolist() == [[51, 55], [51, 75]]
    assert tuple(p.obs.shape) == (50, 40)
    assert np.array_equal(p.obs.numpy()[:, 20:], obs[1][..., 0])
    assert float(p.d[0, 0]) > 0 and float(p.d[60, 70]) == 0.0  # PML damping
    cfg["equation"] = "acoustic_habc"
    q = fwi.problem_from_config(cfg, device="cpu", shots=[1])
    assert q.equation == "acoustic_habc" and q.B == 1 and tuple(q.obs.shape) == (50, 20)


def test_problem_from_config_rejects(tmp_path):
    cfg, _ = _write_case(tmp_path)
    cfg["geom"]["boundary"]["width"] = 40
    with pytest.raises(ValueError, match="width"):
        fwi.problem_from_config(cfg, device="cpu")
    cfg["geom"]["boundary"]["width"] = 50
    cfg["geom"]["nt"] = 60
    with pytest.raises(ValueError, match="records|wavelet"):
        fwi.problem_from_config(cfg, device="cpu")
    cfg["geom"]["nt"] = 50
    cfg["equation"] = "viscoelastic"
    with pytest.raises(ValueError, match="equation"):
        fwi.problem_from_config(cfg, device="cpu")
    cfg["equation"] = "elastic"
    with pytest.raises(ValueError, match="initPath.vs"):
        fwi.problem_from_config(cfg, device="cpu")


def test_elastic_problem_from_config(tmp_path):
    cfg, _ = _write_case(tmp_path)
    nz, nx, nt = 30, 40, 50
    for k, v in (("vs", 1200.0), ("rho", 2000.0)):
        np.save(tmp_path / f"init_{k}.npy", np.full((nz, nx), v))
        cfg["geom"]["initPath"][k] = str(tmp_path / f"init_{k}.npy")
    cfg["equation"] = "elastic"
    cfg["geom"]["invlist"] = {"vp": True, "vs": False, "rho": True}
    cfg["geom"]["source_type"] = ["vz"]
    cfg["geom"]["receiver_type"] = ["vx", "vz"]
    obs = np.empty(2, dtype=np.ndarray)
    for i in range(2):
        obs[i] = np.full((nt, nx // 2, 2), float(i), np.float32)
    np.save(tmp_path / "obs.npy", obs)
    p = fwi.problem_from_config(cfg, device="cpu")
    assert isinstance(p, fwi.ElasticProblem)
    assert p.invlist == ("vp", "rho") and p.src_type == ("vz",) and p.rec_type == ("vx", "vz")
    assert tuple(p.obs.shape) == (nt, nx, 2) and float(p.obs[0, nx // 2, 1]) == 1.0
    assert float(p.models_init["vs"][60, 70]) == 1200.0
    assert float(p.models_true["vs"][60, 70]) == 1200.0  # no truePath.vs: the initial model
