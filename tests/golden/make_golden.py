"""Write golden vectors from the reference's own step functions.

Runs in the build container only: it reads the reference tree under
/root/reference/seistorch, which is absent on a GPU box.  The tests read only
the .npz files this script writes next to it; nothing here is imported by a
test, by smoke() or by bench.py.

It loads single files of the reference by path, never the ``seistorch``
package (its __init__ imports mpi4py, segyio, obspy, h5py, torchaudio, none of
which is installed).  The parent packages are empty stand-ins, and while the
files execute, ``Tensor.to`` sends the import-time "cuda" placement of the
stencil kernels (equations2d/acoustic.py:56-57, acoustic_habc.py:56-57,
convkernel.py:69-91) to the CPU.  For the FP64 vectors the module-level
stencil kernels are replaced by their float64 copies (same values: 0, +-1, -4).

Functions whose vectors are written (reference file:line):
* equations2d/acoustic_habc.py:206-221 ``_time_step`` with ``habc()``
  (:147-204) and ``_habc`` (:79-101); masks and weights from
  habc.py:4-81 (``bound_mask(..., return_idx=True)``,
  ``generate_habc_coefficients_2d``), free surface off and on;
* equations2d/acoustic.py:88-137 and acoustic_habc.py:223-268
  ``_time_step_backward`` and ``_time_step_backward_multiple`` (reverse
  step, ring restore without the top row for a free surface, source re-add);
* equations2d/elastic.py:7-37 ``_time_step``, :39-121 ``_time_step_backward``,
  :123-206 ``_time_step_backward_multiple``;
* equations2d/acoustic_rho_habc.py:31-57 ``_time_step`` (convkernel.py
  face kernels), :59-121 ``_time_step_backward[_multiple]``;
* equations2d/utils.py:50-89 ``save_boundaries`` / ``restore_boundaries``.

The time loops below follow rnn.py:178-212 (step, then the raw source add of
source.py:47-52, then the receiver gather of probe.py:42-44); dt and h are
0-d tensors of the field's dtype, as cell.py:18 and geom.py:32 hold them.

Usage:  python tests/golden/make_golden.py
"""
import importlib.util
import os
import sys
import types

import numpy as np
import torch

REF = "/root/reference/seistorch"
OUT = os.path.dirname(os.path.abspath(__file__))
NZ, NX = 112, 124          # padded grid: 50-cell boundary strips + a 12 x 24 interior
W = 50


# ----------------------------------------------------------------- loading

def _package(name):
    m = types.ModuleType(name)
    m.__path__ = []
    sys.modules[name] = m
    return m


def _load(name, rel):
    spec = importlib.util.spec_from_file_location(name, os.path.join(REF, rel))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_reference():
    _package("seistorch")
    _package("seistorch.equations2d")
    orig_to = torch.Tensor.to

    def cpu(x):
        if isinstance(x, BaseException) or (isinstance(x, str) and x.startswith("cuda")):
            return "cpu"
        if isinstance(x, torch.device) and x.type == "cuda":
            return "cpu"
        return x

    def to_cpu(self, *a, **k):
        a = tuple(cpu(x) for x in a)
        if "device" in k:
            k["device"] = cpu(k["device"])
        return orig_to(self, *a, **k)

    torch.Tensor.to = to_cpu
    try:
        mods = dict(
            utils=_load("seistorch.equations2d.utils", "equations2d/utils.py"),
            convkernel=_load("seistorch.equations2d.convkernel", "equations2d/convkernel.py"),
            acoustic=_load("seistorch.equations2d.acoustic", "equations2d/acoustic.py"),
            acoustic_habc=_load("seistorch.equations2d.acoustic_habc",
                                "equations2d/acoustic_habc.py"),
            elastic=_load("seistorch.equations2d.elastic", "equations2d/elastic.py"),
            rho=_load("seistorch.equations2d.acoustic_rho_habc",
                      "equations2d/acoustic_rho_habc.py"),
            habc=_load("seistorch.habc", "habc.py"),
        )
    finally:
        torch.Tensor.to = orig_to
    return mods


def set_dtype(mods, dtype):
    """Stencil kernels of the loaded modules in `dtype` (values unchanged)."""
    for key in ("acoustic", "acoustic_habc"):
        m = mods[key]
        m.kernel = m.kernel.to(dtype)
    r = mods["rho"]
    for name in ("kernelx_gc_pos", "kernelx_gc_neg", "kernely_gc_pos", "kernely_gc_neg",
                 "kernelx_gc_pos2", "kernelx_gc_neg2", "kernely_gc_pos2", "kernely_gc_neg2",
                 "kernel", "kernelx", "kernely", "kernelx_nc", "kernely_nc"):
        setattr(r, name, getattr(r, name).to(dtype))


# ----------------------------------------------------------------- inputs

def smooth_model(rng, lo, hi, nz=NZ, nx=NX):
    """Layered model with a smooth seeded perturbation, float32."""
    z = np.linspace(0.0, 1.0, nz)[:, None]
    base = lo + (hi - lo) * z * np.ones((1, nx))
    pert = rng.standard_normal((nz, nx))
    for _ in range(6):
        pert = 0.25 * (np.roll(pert, 1, 0) + np.roll(pert, -1, 0)
                       + np.roll(pert, 1, 1) + np.roll(pert, -1, 1))
    pert /= np.abs(pert).max()
    return (base * (1.0 + 0.05 * pert)).astype(np.float32)


def smooth_field(rng, shape):
    f = rng.standard_normal(shape)
    for _ in range(2):
        f = 0.5 * f + 0.125 * (np.roll(f, 1, -1) + np.roll(f, -1, -1)
                               + np.roll(f, 1, -2) + np.roll(f, -1, -2))
    return f.astype(np.float32)


def ricker(fm, dt, nt, delay):
    t = (np.arange(nt) - delay) * dt
    a = (np.pi * fm * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def pml(nz, nx, multiple=False, N=W):
    """Damping of pml.py:29-59 (an input here: its values are pinned by
    closed forms elsewhere), float32."""
    d0 = (1.5 * 1500.0 / N) * np.log10(1e4)
    prof = (d0 * np.linspace(0.0, 1.0, N + 1) ** 2)[::-1]
    pz = np.zeros(nz)
    px = np.zeros(nx)
    if not multiple:
        pz[:N + 1] = prof
    pz[nz - N - 1:] = prof[::-1]
    px[:N + 1] = prof
    px[nx - N - 1:] = prof[::-1]
    return np.sqrt(pz[:, None] ** 2 + px[None, :] ** 2).astype(np.float32)


def t(x, dtype):
    return torch.as_tensor(np.asarray(x)).to(dtype)


def scalar(x, dtype):
    return torch.tensor(float(np.float32(x)), dtype=dtype)


def source_fn(smask):
    """WaveSource.forward2d of source.py:47-57 without source encoding:
    clone (second-order equations) and add smask * X."""
    def f(Y, X, dt=1.0):
        Y_new = Y.clone()
        Y_new += smask * X
        return Y_new
    return f


# ----------------------------------------------------------------- vectors

def gen_habc(mods, rng):
    ah, hb = mods["acoustic_habc"], mods["habc"]
    B = 2
    vp = smooth_model(rng, 1500.0, 2500.0)
    h1 = smooth_field(rng, (B, NZ, NX))
    h2 = smooth_field(rng, (B, NZ, NX))
    dt, h = 1e-3, 10.0
    out = dict(vp=vp, h1=h1, h2=h2, dt=np.float32(dt), h=np.float32(h))
    for mult in (0, 1):
        b = hb.generate_habc_coefficients_2d((NZ, NX), W, multiple=bool(mult))
        tm, bm, lm, rm = hb.bound_mask(NZ, NX, W, "cpu", 1, return_idx=True,
                                       multiple=bool(mult))
        out[f"b_m{mult}"] = b.numpy()
        out[f"mask_bottom_m{mult}"] = bm.numpy().astype(np.uint8)
        out[f"mask_left_m{mult}"] = lm.numpy().astype(np.uint8)
        out[f"mask_right_m{mult}"] = rm.numpy().astype(np.uint8)
        if tm is not None:
            out[f"mask_top_m{mult}"] = tm.numpy().astype(np.uint8)
        masks = hb.bound_mask(NZ, NX, W, "cpu", B, return_idx=True, multiple=bool(mult))
        for dtype, tag in ((torch.float32, "32"), (torch.float64, "64")):
            set_dtype(mods, dtype)
            y, _ = ah._time_step(t(vp, dtype), t(h1, dtype), t(h2, dtype), scalar(dt, dtype),
                                 scalar(h, dtype), b.to(dtype), habcs=masks)
            out[f"y{tag}_m{mult}"] = y.numpy()
    return out


def _loop_acoustic_like(step, params, geoms, nt, w, src, rec_z, B, dtype, kw):
    smask = torch.zeros((B, NZ, NX), dtype=dtype)
    for k, (z, x) in enumerate(src):
        smask[k, z, x] = 1.0
    h1 = torch.zeros((B, NZ, NX), dtype=dtype)
    h2 = torch.zeros_like(h1)
    wt = torch.as_tensor(w).to(dtype)
    recs = []
    for n in range(nt):
        h1, h2 = step(*params, h1, h2, *geoms, **kw)
        y = h1.clone()
        y += smask * wt[n]
        h1 = y
        recs.append(h1[:, rec_z, :].clone())
    return torch.stack(recs, 0).numpy(), h1.numpy()


def gen_habc_loop(mods, rng):
    ah, hb = mods["acoustic_habc"], mods["habc"]
    B, nt, dt, h = 2, 120, 1e-3, 10.0
    vp = smooth_model(rng, 1500.0, 2500.0)
    w = ricker(25.0, dt, nt, 30)
    out = dict(vp=vp, w=w, dt=np.float32(dt), h=np.float32(h), nt=nt)
    for mult in (0, 1):
        # shots in a corner strip and near a side strip; free surface: source at depth 2
        src = [(45, 20), (70, 100)] if not mult else [(2, 30), (40, 95)]
        rec_z = 47 if not mult else 1
        out[f"src_m{mult}"] = np.array(src, np.int64)
        out[f"rec_z_m{mult}"] = rec_z
        b = hb.generate_habc_coefficients_2d((NZ, NX), W, multiple=bool(mult))
        masks = hb.bound_mask(NZ, NX, W, "cpu", B, return_idx=True, multiple=bool(mult))
        for dtype, tag in ((torch.float32, "32"), (torch.float64, "64")):
            set_dtype(mods, dtype)
            recs, last = _loop_acoustic_like(ah._time_step, (t(vp, dtype),),
                                             (scalar(dt, dtype), scalar(h, dtype), b.to(dtype)),
                                             nt, w, src, rec_z, B, dtype, dict(habcs=masks))
            out[f"rec{tag}_m{mult}"] = recs
            if tag == "64":
                out[f"last64_m{mult}"] = last.astype(np.float32)
    return out


def gen_rho(mods, rng):
    rm, hb = mods["rho"], mods["habc"]
    B = 1
    vp = smooth_model(rng, 1500.0, 2500.0)
    rho = smooth_model(rng, 1000.0, 1200.0)
    rho[60:] *= 1.8                                  # density jump
    h1 = smooth_field(rng, (B, NZ, NX))
    h2 = smooth_field(rng, (B, NZ, NX))
    dt, h = 1e-3, 10.0
    out = dict(vp=vp, rho=rho, h1=h1, h2=h2, dt=np.float32(dt), h=np.float32(h))
    for mult in (0, 1):
        b = hb.generate_habc_coefficients_2d((NZ, NX), W, multiple=bool(mult))
        masks = hb.bound_mask(NZ, NX, W, "cpu", B, return_idx=True, multiple=bool(mult))
        for dtype, tag in ((torch.float32, "32"), (torch.float64, "64")):
            set_dtype(mods, dtype)
            y, _ = rm._time_step(t(vp, dtype), t(rho, dtype), t(h1, dtype), t(h2, dtype),
                                 scalar(dt, dtype), scalar(h, dtype), b.to(dtype), habcs=masks)
            out[f"y{tag}_m{mult}"] = y.numpy()
    # records of a short loop through the jump
    nt, B2 = 120, 2
    w = ricker(25.0, dt, nt, 30)
    src = [(55, 30), (75, 90)]
    out.update(loop_w=w, loop_src=np.array(src, np.int64), loop_rec_z=58, loop_nt=nt)
    b = hb.generate_habc_coefficients_2d((NZ, NX), W)
    masks = hb.bound_mask(NZ, NX, W, "cpu", B2, return_idx=True)
    for dtype, tag in ((torch.float32, "32"), (torch.float64, "64")):
        set_dtype(mods, dtype)
        recs, last = _loop_acoustic_like(rm._time_step, (t(vp, dtype), t(rho, dtype)),
                                         (scalar(dt, dtype), scalar(h, dtype), b.to(dtype)),
                                         nt, w, src, 58, B2, dtype, dict(habcs=masks))
        out[f"loop_rec{tag}"] = recs
        if tag == "64":
            out["loop_last64"] = last.astype(np.float32)
    return out


def gen_elastic(mods, rng):
    el = mods["elastic"]
    B = 1
    vp = smooth_model(rng, 1800.0, 3500.0)
    vs = (vp / 1.73).astype(np.float32)
    vs[:58] = 0.0                                     # water layer
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    d = pml(NZ, NX)
    dt, h = 5e-4, 5.0
    f = [smooth_field(rng, (B, NZ, NX)) * s for s in (1e-3, 1e-3, 1e3, 1e3, 1e3)]
    out = dict(vp=vp, vs=vs, rho=rho, d=d, dt=np.float32(dt), h=np.float32(h))
    names = ("vx", "vz", "txx", "tzz", "txz")
    for n, a in zip(names, f):
        out[n] = a
    for dtype, tag in ((torch.float32, "32"), (torch.float64, "64")):
        res = el._time_step(t(vp, dtype), t(vs, dtype), t(rho, dtype), *(t(a, dtype) for a in f),
                            scalar(dt, dtype), scalar(h, dtype), t(d, dtype))
        for n, a in zip(names, res):
            out[f"{n}{tag}"] = a.numpy()
    # loop: explosive source on txx and tzz, vx and vz recorded on a row
    nt, B2 = 160, 2
    w = ricker(30.0, dt, nt, 40)
    src = [(56, 40), (70, 85)]
    rec_z = 52
    out.update(loop_w=w, loop_src=np.array(src, np.int64), loop_rec_z=rec_z, loop_nt=nt)
    for dtype, tag in ((torch.float32, "32"), (torch.float64, "64")):
        smask = torch.zeros((B2, NZ, NX), dtype=dtype)
        for k, (z, x) in enumerate(src):
            smask[k, z, x] = 1.0
        fld = [torch.zeros((B2, NZ, NX), dtype=dtype) for _ in names]
        wt = torch.as_tensor(w).to(dtype)
        recs = []
        args = (t(vp, dtype), t(vs, dtype), t(rho, dtype))
        for n in range(nt):
            fld = list(el._time_step(*args, *fld, scalar(dt, dtype), scalar(h, dtype),
                                     t(d, dtype)))
            for k in (2, 3):                          # txx, tzz: in-place add (first order)
                fld[k] += smask * wt[n]
            recs.append(torch.stack([fld[0][:, rec_z, :], fld[1][:, rec_z, :]], -1).clone())
        out[f"loop_rec{tag}"] = torch.stack(recs, 0).numpy()
        if tag == "64":
            out["loop_last64"] = np.stack([a.numpy() for a in fld]).astype(np.float32)
    return out


def gen_reverse(mods, rng):
    """Reverse steps of the boundary-saving protocol (one call each), FP32."""
    ac, ah, el, rm, ut, hb = (mods[k] for k in ("acoustic", "acoustic_habc", "elastic", "rho",
                                                "utils", "habc"))
    set_dtype(mods, torch.float32)
    B = 2
    vp = smooth_model(rng, 1500.0, 2500.0)
    rho = smooth_model(rng, 1000.0, 2000.0)
    h1 = smooth_field(rng, (B, NZ, NX))
    h2 = smooth_field(rng, (B, NZ, NX))
    ring_src = smooth_field(rng, (B, NZ, NX))
    src = [(60, 30), (52, 80)]
    smask = np.zeros((B, NZ, NX), np.float32)
    for k, (z, x) in enumerate(src):
        smask[k, z, x] = 1.0
    wv = np.float32(0.75)
    dt, h = 1e-3, 10.0
    out = dict(vp=vp, rho=rho, h1=h1, h2=h2, ring_src=ring_src, smask=smask, wv=wv,
               dt=np.float32(dt), h=np.float32(h))
    f32 = torch.float32
    rs = t(ring_src, f32)
    srcf = source_fn(t(smask, f32))
    # rings: save_boundaries of a field, and the restore into a copy of another
    for mult in (0, 1):
        ring = ut.save_boundaries(rs)
        for name, r in zip(("top", "bottom", "left", "right"), ring):
            out[f"ring_{name}"] = r.numpy()
        restored = ut.restore_boundaries(t(h1, f32).clone(), ring, multiple=bool(mult))
        out[f"restored_m{mult}"] = restored.numpy()

    def rings():
        return [ut.save_boundaries(rs), ut.save_boundaries(rs)]

    dpml = {m: pml(NZ, NX, bool(m)) for m in (0, 1)}
    out["d_m0"], out["d_m1"] = dpml[0], dpml[1]
    geoms = (scalar(dt, f32), scalar(h, f32))
    y, _ = ac._time_step_backward(t(vp, f32), t(h1, f32), t(h2, f32), *geoms, t(dpml[0], f32),
                                  rings(), [["h1"], srcf, t(wv, f32)])
    out["acoustic_back"] = y.numpy()
    y, _ = ac._time_step_backward_multiple(t(vp, f32), t(h1, f32), t(h2, f32), *geoms,
                                           t(dpml[1], f32), rings(), [["h1"], srcf, t(wv, f32)])
    out["acoustic_back_multiple"] = y.numpy()
    b0 = hb.generate_habc_coefficients_2d((NZ, NX), W)
    b1 = hb.generate_habc_coefficients_2d((NZ, NX), W, multiple=True)
    y, _ = ah._time_step_backward(t(vp, f32), t(h1, f32), t(h2, f32), *geoms, b0, rings(),
                                  [["h1"], srcf, t(wv, f32)])
    out["habc_back"] = y.numpy()
    y, _ = ah._time_step_backward_multiple(t(vp, f32), t(h1, f32), t(h2, f32), *geoms, b1,
                                           rings(), [["h1"], srcf, t(wv, f32)])
    out["habc_back_multiple"] = y.numpy()
    y, _ = rm._time_step_backward(t(vp, f32), t(rho, f32), t(h1, f32), t(h2, f32), *geoms, b0,
                                  rings(), [["h1"], srcf, t(wv, f32)])
    out["rho_back"] = y.numpy()
    r = rm._time_step_backward_multiple(t(vp, f32), t(rho, f32), t(h1, f32), t(h2, f32), *geoms,
                                        b1, rings(), [["h1"], srcf, t(wv, f32)])
    out["rho_back_multiple"] = r[0].numpy()

    # elastic reverse steps: five fields, rings of five fields, explosive source
    vpe = smooth_model(rng, 1800.0, 3500.0)
    vse = (vpe / 1.73).astype(np.float32)
    vse[:58] = 0.0
    rhoe = (310.0 * vpe.astype(np.float64) ** 0.25).astype(np.float32)
    dte, he = 5e-4, 5.0
    fe = [smooth_field(rng, (1, NZ, NX)) * s for s in (1e-3, 1e-3, 1e3, 1e3, 1e3)]
    re = [smooth_field(rng, (1, NZ, NX)) * s for s in (1e-3, 1e-3, 1e3, 1e3, 1e3)]
    names = ("vx", "vz", "txx", "tzz", "txz")
    out.update(el_vp=vpe, el_vs=vse, el_rho=rhoe, el_dt=np.float32(dte), el_h=np.float32(he))
    for n, a, r_ in zip(names, fe, re):
        out[f"el_{n}"] = a
        out[f"el_ring_{n}"] = r_
    wve = np.float32(2.5)
    out["el_wv"] = wve
    for mult in (0, 1):
        ringe = [ut.save_boundaries(t(r_, f32)) for r_ in re]

        def srce(Y, X, dt_=1.0, _m=t(smask[:1], f32)):
            Y_new = Y                                    # first-order: in place
            Y_new += _m * X
            return Y_new
        fn = el._time_step_backward_multiple if mult else el._time_step_backward
        res = fn(t(vpe, f32), t(vse, f32), t(rhoe, f32), *(t(a, f32) for a in fe),
                 scalar(dte, f32), scalar(he, f32), t(dpml[mult], f32), ringe,
                 [["txx", "tzz"], srce, t(wve, f32)])
        for n, a in zip(names, res):
            out[f"el_back_m{mult}_{n}"] = a.detach().numpy()
    return out


def main():
    if not os.path.isdir(REF):
        raise SystemExit(f"{REF} not found: golden vectors are written in the build container only")
    torch.set_num_threads(4)
    mods = load_reference()
    rng = np.random.default_rng(20230503)
    for name, fn in (("habc_step", gen_habc), ("habc_loop", gen_habc_loop), ("rho", gen_rho),
                     ("elastic", gen_elastic), ("reverse", gen_reverse)):
        out = fn(mods, rng)
        path = os.path.join(OUT, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


if __name__ == "__main__":
    main()
