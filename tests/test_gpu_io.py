"""A configuration's files through the HIP forward: model .npy, geometry
.pkl -> padded acquisition -> propagate -> per-shot records written as the
reference's .npy object array -> read back, against the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import acoustic as oa

pytestmark = pytest.mark.gpu


def test_files_to_records_round_trip(tmp_path):
    from seistorch_amd import io as sio
    from seistorch_amd import synthetic
    from seistorch_amd.equations2d import acoustic as ac
    from seistorch_amd.pml import generate_pml_coefficients_2d
    from seistorch_amd.utils import pad_model, ricker_wave
    nz_i, nx_i, bw, nt, dt, h = 40, 64, 50, 300, 1e-3, 10.0
    v = synthetic.layered_model(nz_i, nx_i, 1500.0, 3000.0, seed=3)
    np.save(tmp_path / "vp.npy", v)
    sources = [[8 + 16 * i, 1] for i in range(3)]
    receivers = [[list(range(nx_i)), [5] * nx_i] for _ in range(3)]  # fixed, as the examples
    receivers[2] = [list(range(0, nx_i, 3)), [4] * len(range(0, nx_i, 3))]
    sio.write_pkl(tmp_path / "sources.pkl", sources)
    sio.write_pkl(tmp_path / "receivers.pkl", receivers)
    cfg = {"geom": {"sources": str(tmp_path / "sources.pkl"),
                    "receivers": str(tmp_path / "receivers.pkl"),
                    "obsPath": str(tmp_path / "obs.npy"),
                    "boundary": {"width": bw}, "multiple": False}}
    vp = pad_model(sio.load_model(str(tmp_path / "vp.npy"))).astype(np.float32)
    nz, nx = vp.shape
    src, rec = sio.acquisition_from_cfg(cfg)
    d = generate_pml_coefficients_2d((nz, nx), bw).numpy()
    w = ricker_wave(15.0, dt, nt, delay=60).numpy()
    flat = synthetic.rec_flat_index(rec, nz, nx)
    out = ac.propagate(torch.from_numpy(vp).cuda(), torch.from_numpy(d).cuda(),
                       torch.from_numpy(w).cuda(), torch.from_numpy(src).cuda(),
                       torch.from_numpy(flat).cuda(), len(sources), dt, h)
    rs = sio.SeisRecord(cfg)
    rs.setup("forward")
    shots = sio.split_records(out.cpu().numpy(), [len(r) for r in rec])
    for i, shot in enumerate(shots):
        rs[i] = shot
    rs.write()
    back = sio.SeisRecord(cfg)
    back.setup("inversion")
    ref = oa.forward(vp, d, dt, h, w, src, rec)
    for i in range(3):
        assert back[i].shape == (nt, len(receivers[i][0]), 1)
        assert np.array_equal(back[i], shots[i])  # the files hold exactly what was written
        # the padded coordinates reach the oracle's cells: a shifted source or
        # receiver would differ at O(1); FP32 propagation leaves ~1e-5
        e = np.linalg.norm(back[i] - ref[i]) / np.linalg.norm(ref[i])
        print(f"shot {i}: {back[i].shape} records vs FP64 oracle {e:.2e}")
        assert e < 1e-4
