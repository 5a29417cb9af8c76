"""Benchmark: acoustic FWI gradient (forward + hand-written adjoint) on one MI355X.

Workload (one "step"): the full L2-misfit gradient of 64 shots on an
800 x 2400 model at dx = 10 m (900 x 2500 with the 50-cell PML), 8 Hz Ricker,
dt = 0.8 ms, nt = 5000, 2400 surface receivers: forward loop with boundary
saving, misfit, adjoint loop with reconstruction and gradient accumulation.
Inputs (model, observed records) are resident in HBM before timing starts.

Metric: Gcell-updates/s = shots * nz_pad * nx_pad * nt / wall-time of the step
(each cell, step and shot counted once for forward+adjoint together).

Prints ONE JSON line.  ``--gpus N`` with N > 1 exits: this build measures one
GPU only.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "Gcell-updates/s fwd+adjoint, % HBM roofline; wall-clock to target misfit"
HELD_OUT_SEED = 20240611  # generator stream of the preconditioner's held-out pairs
HBM_PEAK = 8.0e12  # B/s, MI355X
# algorithmic bytes per cell-update of each timed kernel (fields touched once
# per step, damping/coefficients read once per cell)
# (k_fwd2 advances two steps per launch)
BYTES_PER_CELL = {"k_fwd": 16.0, "k_bwd": 36.0, "k_fwd2": 32.0, "k_bwd2": 72.0,
                  "k_habc_fwd": 16.0, "k_habc_adj": 36.0,
                  "k_el_fwd": 52.0, "k_el_bwd": 116.0,
                  "k_fwd_batch": 16.0, "k_fwd2_batch": 32.0, "k_bwd_batch": 36.0}
KERNEL_IDS = {"acoustic": (("k_fwd", 0), ("k_bwd", 1), ("k_fwd2", 8), ("k_bwd2", 9)),
              "habc": (("k_habc_fwd", 4), ("k_habc_adj", 5)),
              "elastic": (("k_el_fwd", 2), ("k_el_bwd", 3)),
              "generator": (("k_fwd2_batch", 6), ("k_bwd_batch", 7))}  # slot 6: the batched forward, two steps per launch


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-steps", type=int, default=240,
                   help="time steps of the bounded CPU sample (1 shot, full grid)")
    p.add_argument("--nt", type=int, default=None,
                   help="shorter time axis for profiler runs (the metric is quoted at full nt)")
    p.add_argument("--workload", choices=["acoustic", "habc", "elastic", "generator", "precond",
                                          "filter", "postprocess", "rho"],
                   default="acoustic",
                   help="acoustic: 64-shot 800x2400 FWI gradient (the headline); habc: the "
                        "same with the HABC boundary; elastic: "
                        "32-shot 600x1800 vp/vs/rho gradient; generator: 256 uneven seeded "
                        "models, records + misfit gradient in one batch; precond: train the "
                        "U-Net preconditioner on generator pairs, then plain vs preconditioned "
                        "FWI on the 800x2400 problem; filter: zero-phase Butterworth filter "
                        "of the 64-shot 800x2400 problem's records (multiscale FWI); "
                        "postprocess: smoothing, seabed cut and illumination of one "
                        "900x2500 gradient; rho: variable-density HABC forward records of "
                        "the 64-shot 800x2400 problem")
    p.add_argument("--models", type=int, default=256, help="generator: number of models")
    p.add_argument("--graph", type=int, default=0,
                   help="generator: 1 records each batched loop into one hipGraph (kernel "
                        "timing events are then unavailable)")
    p.add_argument("--snapshot-every", type=int, default=500,
                   help="generator: field-pair snapshot spacing of the batched adjoint")
    p.add_argument("--pair-models", type=int, default=5000,
                   help="precond: generator models whose gradients make the training pairs")
    p.add_argument("--train-steps", type=int, default=5000, help="precond: AdamW steps")
    p.add_argument("--pair-first", type=int, default=0,
                   help="precond: index of the first generator model of this run's pairs")
    p.add_argument("--val-models", type=int, default=250,
                   help="precond: held-out models (a second seeded stream)")
    p.add_argument("--init-net", default=None,
                   help="precond: continue training this saved network (its .json is carried)")
    p.add_argument("--save-net", default=None,
                   help="precond: save the trained network (torch state dict + .json)")
    p.add_argument("--load-net", default=None,
                   help="precond: FWI with a saved network instead of training one")
    p.add_argument("--precond-model", choices=["generator", "anomaly"], default="generator",
                   help="precond: true model of the FWI comparison (the generator's family "
                        "at 800x2400, or the anomaly model of the FWI-gradient workload)")
    p.add_argument("--fwi-iters", type=int, default=30, help="precond: FWI iterations per method")
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run a plain 1 GiB device copy (HBM-counter calibration)")
    return p.parse_args()


def build_problem(dev, seed=20230503, nt_override=None):
    import torch
    from seistorch_amd import synthetic
    from seistorch_amd.pml import generate_pml_coefficients_2d
    from seistorch_amd.utils import pad_model, ricker_wave
    nz_i, nx_i, nshots, nt, dt, h, fm = 800, 2400, 64, 5000, 0.8e-3, 10.0, 8.0
    nt = nt_override or nt
    true, init = synthetic.anomaly_model(nz_i, nx_i, seed=seed)
    vt = pad_model(true)
    vi = pad_model(init)
    nz, nx = vi.shape
    src, rec = synthetic.surface_acquisition(nx_i, nshots, src_depth=2)
    P = dict(nz=nz, nx=nx, B=nshots, nt=nt, dt=dt, h=h, fm=fm, nz_i=nz_i, nx_i=nx_i)
    P["vp_true"] = torch.from_numpy(vt).to(dev)
    P["vp_init"] = torch.from_numpy(vi).to(dev)
    P["d"] = generate_pml_coefficients_2d((nz, nx), 50).to(dev)
    P["w"] = ricker_wave(fm, dt, nt, delay=256).to(dev)
    P["src"] = torch.from_numpy(src.astype(np.int32)).to(dev)
    P["rec"] = torch.from_numpy(synthetic.rec_flat_index(rec, nz, nx)).to(dev)
    P["src_np"], P["rec_np"] = src, rec
    P["vp_init_np"], P["d_np"], P["w_np"] = vi, P["d"].cpu().numpy(), P["w"].cpu().numpy()
    return P


def build_elastic_problem(dev, seed=20230503, nt_override=None):
    import torch
    from seistorch_amd import synthetic
    from seistorch_amd.pml import generate_pml_coefficients_2d
    from seistorch_amd.utils import pad_model, ricker_wave
    nz_i, nx_i, nshots, nt, dt, h, fm = 600, 1800, 32, 6000, 0.5e-3, 5.0, 10.0
    nt = nt_override or nt
    vp, vs, rho = synthetic.elastic_models(nz_i, nx_i, seed=seed)
    vp0, vs0, rho0 = synthetic.elastic_models(nz_i, nx_i, pert=0.0, seed=seed)
    src, rec = synthetic.surface_acquisition(nx_i, nshots, src_depth=2)
    P = dict(nz=nz_i + 100, nx=nx_i + 100, B=nshots, nt=nt, dt=dt, h=h, fm=fm, nz_i=nz_i,
             nx_i=nx_i, elastic=True)
    for k, a in (("vp_true", vp), ("vs_true", vs), ("rho_true", rho), ("vp_init", vp0),
                 ("vs_init", vs0), ("rho_init", rho0)):
        P[k] = torch.from_numpy(pad_model(a).astype(np.float32)).to(dev)
    P["d"] = generate_pml_coefficients_2d((P["nz"], P["nx"]), 50).to(dev)
    P["w"] = ricker_wave(fm, dt, nt, delay=160).to(dev)
    P["src"] = torch.from_numpy(src.astype(np.int32)).to(dev)
    P["rec"] = torch.from_numpy(synthetic.rec_flat_index(rec, P["nz"], P["nx"])).to(dev)
    return P


def elastic_records(P, vp, vs, rho):
    from seistorch_amd.equations2d import elastic as el
    return el.propagate(vp, vs, rho, P["d"], P["w"], P["src"], P["rec"], P["B"], P["dt"],
                        P["h"], ("txx", "tzz"), ("vx", "vz"))


def elastic_gradient_step(P):
    import torch
    m = [P[k].detach().requires_grad_(True) for k in ("vp_init", "vs_init", "rho_init")]
    rec = elastic_records(P, *m)
    loss = torch.sum((rec - P["obs"]) ** 2)
    loss.backward()
    return loss.detach(), m[0].grad


def habc_records(P, vp):
    from seistorch_amd.equations2d import acoustic_habc as ah
    return ah.propagate(vp, P["b"], P["w"], P["src"], P["rec"], P["B"], P["dt"], P["h"])


def habc_gradient_step(P):
    import torch
    vp = P["vp_init"].detach().requires_grad_(True)
    rec = habc_records(P, vp)
    loss = torch.sum((rec - P["obs"]) ** 2)
    loss.backward()
    return loss.detach(), vp.grad


def gradient_step(P):
    import torch
    from seistorch_amd.equations2d import acoustic as ac
    vp = P["vp_init"].detach().requires_grad_(True)
    rec = ac.propagate(vp, P["d"], P["w"], P["src"], P["rec"], P["B"], P["dt"], P["h"])
    loss = torch.sum((rec - P["obs"]) ** 2)
    loss.backward()
    return loss.detach(), vp.grad


def _launch_shape_json(workload, shape, tune):
    """The tuned launch shape and the candidates' times for the JSON line."""
    if workload == "elastic":
        out = {"shots_per_group_forward": shape[0], "shots_per_group_adjoint": shape[1]}
        times = {f"{k[0]}_S{k[1]}": round(t, 2) for k, t in tune.items()}
    elif len(shape) == 3:
        out = {"shots_per_group": shape[0], "rows_per_thread": shape[1], "vec_x": shape[2]}
        times = {f"S{a}_rz{b}_vx{c}": round(t, 2) for (a, b, c), t in tune.items()}
    else:
        out = {"shots_per_group": shape[0]}
        times = {f"S{k[0]}": round(t, 2) for k, t in tune.items()}
    out["autotune_ms"] = times
    return out


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(P, steps):
    """The FP32 torch restatement of the reference path on the host cores:
    1 shot on the full padded grid for `steps` time steps (forward with ring
    saving + per-step recompute backward), scaled per cell-update."""
    import torch
    from oracle import ref_torch
    threads = torch.get_num_threads()  # honours OMP_NUM_THREADS (the host's CPU share)
    w = P["w_np"][:steps]
    src = P["src_np"][:1]
    rec = P["rec_np"][:1]
    obs = [torch.zeros((steps, len(rec[0]), 1))]
    t0 = time.perf_counter()
    ref_torch.gradient_bs(P["vp_init_np"], P["d_np"], P["dt"], P["h"], w, src, rec, obs)
    el = time.perf_counter() - t0
    cells = P["nz"] * P["nx"] * steps
    return {"value": cells / el / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "cpu": cpu_model(), "kind": "port",
            "sample": f"1 shot, {P['nz']}x{P['nx']} padded grid, {steps} of {P['nt']} steps, "
                      f"fwd+adjoint (oracle/ref_torch.gradient_bs, torch {torch.__version__}), "
                      f"{el:.1f} s"}


def cpu_baseline_ad(P, workload):
    """The reference's autograd path (boundary_saving off: autograd through
    the loop of the FP32 torch restatement) on the host cores for the HABC
    and elastic workloads: 1 shot on the full padded grid over the first
    steps of the time axis (16 HABC steps, 24 elastic steps: 10-30 s),
    forward + backward, scaled per cell-update."""
    import torch
    from oracle import ref_torch
    from seistorch_amd import synthetic
    threads = torch.get_num_threads()
    src, rec = synthetic.surface_acquisition(P["nx_i"], P["B"], src_depth=2)
    w = P["w"].cpu().numpy()
    if workload == "habc":
        steps, fn = 16, "oracle/ref_torch.habc_gradient_ad"
        c, b = P["vp_init"].cpu().numpy(), P["b"].cpu().numpy()

        def run(n):
            G = np.ones((n, len(rec[0])), np.float32)
            ref_torch.habc_gradient_ad(c, b, P["dt"], P["h"], w[:n], src[:1], rec[:1], G)
    else:
        steps, fn = 24, "oracle/ref_torch.elastic_gradient_ad"
        models = [P[k].cpu().numpy() for k in ("vp_init", "vs_init", "rho_init", "d")]

        def run(n):
            G = np.ones((n, len(rec[0]), 2), np.float32)
            ref_torch.elastic_gradient_ad(*models, P["dt"], P["h"], w[:n], src[:1], rec[:1],
                                          None, rec_cotangent=G)
    run(2)  # first-touch of the autograd buffers, untimed
    t0 = time.perf_counter()
    run(steps)
    el = time.perf_counter() - t0
    cells = P["nz"] * P["nx"] * steps
    return {"value": cells / el / 1e9, "unit": "Gcell-updates/s", "cores": threads,
            "cpu": cpu_model(), "kind": "port",
            "sample": f"1 shot, {P['nz']}x{P['nx']} padded grid, {steps} of {P['nt']} steps, "
                      f"fwd+adjoint by autograd ({fn}, torch {torch.__version__}), {el:.1f} s"}


def copy_calibration(dev, reps=5):
    """The library's float4 streaming copy of 1 GiB (seis_stream_copy),
    timed with HIP events on its stream: its rate is the practical HBM
    ceiling, and its FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc (1 GiB each
    way per dispatch) calibrate the counters of the other kernels.  Returns
    {"bytes_each_way", "tb_per_s"} (read + write bytes over the best time)."""
    import ctypes
    import torch
    from seistorch_amd import _lib
    n = 1 << 28  # 1 GiB of float32
    a = torch.ones(n, device=dev)
    b = torch.empty_like(a)
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    lib = _lib.load()
    rates = {}
    for per_cu in (2, 4, 8, 16, 32):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            lib.seis_profile_enable(1)
            _lib.call("seis_stream_copy", _lib.ptr(a), _lib.ptr(b), n, per_cu * cus,
                      _lib.stream_of(a))
            lib.seis_profile_enable(0)
            ms = ctypes.c_double(0.0)
            cnt = ctypes.c_int64(0)
            _lib.call("seis_profile_collect", 10, ctypes.byref(ms), ctypes.byref(cnt))
            best = ms.value if best is None else min(best, ms.value)
        rates[per_cu] = round(2 * 4 * n / (best / 1e3) / 1e12, 3)
    torch.cuda.synchronize()
    del a, b
    return {"bytes_each_way": 4 * n, "tb_per_s": max(rates.values()),
            "tb_per_s_by_blocks_per_cu": rates}


def run_generator(args, dev):
    """Uneven batch: 256 seeded models (grids 64..512 x 64..768, 4-16 shots,
    nt = 1.2 x diagonal traveltime).  One step = records of every true model
    (batched forward) + the L2-misfit gradient of every start model (batched
    forward + adjoint).  value = fwd+adjoint cell-updates of the gradient pass
    / its wall time, as for the single-model workloads."""
    import ctypes
    import torch
    from seistorch_amd import _lib, autotune, generator
    t0 = time.perf_counter()
    spec = generator.GenSpec(n_models=args.models)
    host = generator.make_instances(spec=spec)
    if args.nt:
        for it in host:
            it["nt"] = min(it["nt"], args.nt)
            it["wavelet"] = it["wavelet"][:it["nt"]]
    insts = generator.to_device(host, dev)
    cells = generator.cell_updates(host)
    print(f"[bench] generator: {len(host)} models, {cells:.3e} cell-updates per propagation, "
          f"built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    shape, tune = autotune.tune_batch(generator.batch_instances(insts, "vp_start"))
    launch = dict(shots_per_group=shape[0], rows_per_thread=shape[1], vec_x=shape[2])
    print(f"[bench] autotuned batch launch shape {shape}", file=sys.stderr, flush=True)
    use_graph = bool(args.graph)
    lib = _lib.load()

    def step():
        torch.cuda.synchronize()
        a = time.perf_counter()
        obs = generator.observed(insts, use_graph=use_graph, **launch)
        torch.cuda.synchronize()
        b = time.perf_counter()
        loss, grads = generator.misfit_gradients(insts, obs, use_graph=use_graph,
                                                 snapshot_every=args.snapshot_every, **launch)
        torch.cuda.synchronize()
        c = time.perf_counter()
        return b - a, c - b, loss

    for i in range(args.warmup):
        step()
        print(f"[bench] warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    lib.seis_profile_enable(0 if use_graph else 1)
    t_fwd = t_grad = 0.0
    for i in range(args.steps):
        f, g, loss = step()
        t_fwd += f
        t_grad += g
        print(f"[bench] step {i + 1}/{args.steps}: records {f:.2f} s, gradient {g:.2f} s",
              file=sys.stderr, flush=True)
    lib.seis_profile_enable(0)
    kern = {}
    for name, kid in KERNEL_IDS["generator"]:
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_int64(0)
        _lib.call("seis_profile_collect", kid, ctypes.byref(ms), ctypes.byref(cnt))
        kern[name] = (ms.value, cnt.value)
    value = cells * args.steps / t_grad / 1e9
    roof = None
    if not use_graph:
        # per launch the batched kernels cover every running instance; the
        # algorithmic bytes of all timed launches over their summed duration
        tot_ms, n_launch = kern["k_bwd_batch"]
        achieved = BYTES_PER_CELL["k_bwd_batch"] * cells * args.steps / (tot_ms / 1e3)
        # measured HBM bytes per cell-update (rocprof counters) x the average
        # cell-updates of one launch
        traffic = None
        tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                             "traffic.json")
        per_cell = json.load(open(tpath)).get("k_bwd_batch_bytes_per_cell") \
            if os.path.exists(tpath) else None
        if per_cell and n_launch:
            traffic = round(per_cell * cells * args.steps / n_launch)
        roof = {"bound": "hbm", "kernel": "k_bwd_batch", "achieved": round(achieved / 1e9, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                "traffic": traffic, "avg_launch_us": round(tot_ms / max(n_launch, 1) * 1e3, 2),
                "launches": n_launch, "bytes_per_cell_update": BYTES_PER_CELL["k_bwd_batch"]}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round((t_fwd + t_grad) / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded layered/faulted/anomaly models, records from the true models)",
        "config": {"workload": "uneven batch: records + L2-misfit gradient of every model, "
                               "PML 50, one launch per step over all running models",
                   "models": len(host), "cell_updates_per_propagation": cells,
                   "nt_range": [min(i["nt"] for i in host), max(i["nt"] for i in host)],
                   "shots_range": [min(i["B"] for i in host), max(i["B"] for i in host)],
                   "hipgraph": use_graph, "snapshot_every": args.snapshot_every},
        "records_gcells_per_s": round(cells * args.steps / t_fwd / 1e9, 3),
        "gradient_s_per_step": round(t_grad / args.steps, 3),
        "records_s_per_step": round(t_fwd / args.steps, 3),
        "roofline": roof,
        "kernels_ms_total": {k: round(v[0], 2) for k, v in kern.items()},
        "launch_shape": {"shots_per_group": shape[0], "rows_per_thread": shape[1],
                         "vec_x": shape[2],
                         "autotune_ms": {f"S{a}_rz{b}_vx{c}": round(v, 2)
                                         for (a, b, c), v in tune.items()}},
        "loss_sum": float(loss.sum()),
    }
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(dev)
    print(json.dumps(out), flush=True)


def run_filter(args, dev):
    """Multiscale record filter: the zero-phase 1-3 Hz order-3 Butterworth
    band (the reference template's first scale) over the records of the
    64-shot 800x2400 problem (nt = 5000, dt = 0.8 ms, 2400 receivers per shot,
    153600 traces).  One step = one filtfilt of all records (forward and
    backward recursion in FP64, one kernel).  value = samples (nt x traces)
    per second."""
    import torch
    from seistorch_amd import signal as sg
    nt, shots, nrec, dt = 5000, 64, 2400, 0.8e-3
    if args.nt:
        nt = args.nt
    b, a = sg.butter(3, [1.0, 3.0], dt)
    g = torch.Generator(device=dev).manual_seed(20230503)
    x = torch.randn((nt, shots * nrec), device=dev, generator=g)
    samples = x.numel()
    for _ in range(args.warmup):
        sg.filtfilt(x, b, a)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[2 * i].record()
        y = sg.filtfilt(x, b, a)
        ev[2 * i + 1].record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    launch_s = sum(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)) \
        / args.steps / 1e3
    bytes_per_sample = 24.0  # read x (4), write + read the FP64 pass (16), write y (4)
    achieved = bytes_per_sample * samples / launch_s
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if os.path.exists(tpath):
        per = json.load(open(tpath)).get("k_filtfilt_bytes_per_sample")
        traffic = round(per * samples) if per else None
    out = {
        "metric": "Gsamples/s zero-phase record filter (multiscale FWI)",
        "value": round(samples * args.steps / el / 1e9, 3), "unit": "Gsamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Gaussian records of the 64-shot 800x2400 problem's shape)",
        "config": {"workload": "zero-phase Butterworth filter of records (filtfilt, FP64)",
                   "nt": nt, "traces": shots * nrec, "dt": dt, "band_hz": [1.0, 3.0],
                   "order": 3, "coefficients": len(b)},
        "roofline": {"bound": "hbm", "kernel": "k_filtfilt", "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": traffic, "avg_launch_us": round(launch_s * 1e6, 2),
                     "launches": args.steps, "bytes_per_sample": bytes_per_sample},
        "checksum": float(y.double().abs().sum()),
    }
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(dev)
    if not args.no_cpu_baseline:
        from oracle import signal as osg
        xs = x.double().cpu().numpy()  # the whole workload: 64 shots, full nt
        t1 = time.perf_counter()
        osg.filtfilt(b, a, xs)
        cel = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(xs.size / cel / 1e9, 5), "unit": "Gsamples/s",
                               "cores": 1, "cpu": cpu_model(), "kind": "port",
                               "sample": f"all {shots * nrec} traces x {nt} samples, "
                                         f"oracle/signal.filtfilt (numpy FP64, vectorised over "
                                         f"traces, one thread), {cel:.1f} s"}
    print(json.dumps(out), flush=True)


def run_postprocess(args, dev):
    """Gradient post-processing between the adjoint and the update
    (seistorch_dist.py:270-280): seabed pre-mask, 5 Gaussian smoothing
    passes per axis (radius 10, sigma 3: 21 taps), seabed cut, division by the
    illumination, on one 900x2500 gradient (the padded 800x2400 grid).  One
    step = one seis_smooth_gradient call (10 kernel launches).  value = cell
    passes per second."""
    import torch
    from seistorch_amd import process as pp
    nz, nx = 900, 2500
    sm = {"counts": 5, "sigma": {"z": 3.0, "x": 3.0}, "radius": {"z": 10, "x": 10}}
    g = torch.Generator(device=dev).manual_seed(20230504)
    grad = torch.randn((nz, nx), device=dev, generator=g)
    mask = torch.ones((nz, nx), device=dev)
    mask[:150] = 0.0
    illum = torch.rand((nz, nx), device=dev, generator=g) + 0.1
    passes = 2 * sm["counts"]
    for _ in range(args.warmup):
        pp.smooth_gradient(grad, sm, pre_mask=mask, post_mask=mask, illum=illum)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[2 * i].record()
        y = pp.smooth_gradient(grad, sm, pre_mask=mask, post_mask=mask, illum=illum)
        ev[2 * i + 1].record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    call_s = sum(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)) \
        / args.steps / 1e3
    cells = nz * nx
    # per pass: read + write the field (8 B/cell); the first pass also reads
    # the pre-mask, the last the cut mask and the illumination (12 B/cell)
    algo_bytes = cells * (8 * passes + 12)
    achieved = algo_bytes / call_s
    out = {
        "metric": "Gcell-passes/s gradient smoothing + seabed cut + illumination",
        "value": round(cells * passes * args.steps / el / 1e9, 3), "unit": "Gcell-passes/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian gradient, uniform illumination, 150-row water layer)",
        "config": {"workload": "gradient post-processing of one 900x2500 gradient",
                   "counts": sm["counts"], "radius": 10, "sigma": 3.0, "taps": 21},
        "roofline": {"bound": "hbm", "kernel": "k_gauss (10 launches per call)",
                     "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                     "avg_launch_us": round(call_s * 1e6 / passes, 2),
                     "launches": args.steps * passes, "bytes_per_call": algo_bytes},
        "checksum": float(y.double().abs().sum()),
    }
    if not args.no_cpu_baseline:
        from oracle import process as opp
        gs, ms, il = grad.cpu().numpy(), mask.cpu().numpy(), illum.cpu().numpy()
        t1 = time.perf_counter()
        opp.postprocess(gs, sm, pre_mask=ms, post_mask=ms, illum=il)
        cel = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(cells * passes / cel / 1e9, 5),
                               "unit": "Gcell-passes/s", "cores": 1, "cpu": cpu_model(),
                               "kind": "port",
                               "sample": f"the whole call on the same gradient, "
                                         f"oracle/process.postprocess (numpy FP64, one "
                                         f"thread), {cel:.2f} s"}
    print(json.dumps(out), flush=True)


def run_rho(args, dev):
    """Variable-density acoustic equation with the HABC boundary
    (acoustic_rho_habc.py): records of the 64-shot 800x2400 problem (padded
    900x2500) over a density model rising with depth, 1000 time steps (the
    per-step rate does not depend on nt).  One step = one forward loop.
    value = cell-updates per second."""
    import torch
    from seistorch_amd.equations2d import acoustic_rho_habc as ar
    from seistorch_amd.habc import generate_habc_coefficients_2d
    P = build_problem(dev, nt_override=args.nt or 1000)
    nz, nx, B, nt = P["nz"], P["nx"], P["B"], P["nt"]
    z = torch.arange(nz, device=dev, dtype=torch.float32)[:, None]
    rho = (1000.0 + 1.2 * z).expand(nz, nx).contiguous()
    b = generate_habc_coefficients_2d((nz, nx), 50).to(dev)

    def fwd():
        return ar.propagate(P["vp_true"], rho, b, P["w"], P["src"], P["rec"], B, P["dt"], P["h"])

    for _ in range(args.warmup):
        fwd()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[2 * i].record()
        rec = fwd()
        ev[2 * i + 1].record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    loop_s = sum(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)) \
        / args.steps / 1e3
    cells = B * nz * nx * nt
    # per cell-update: read u^n and u^{n-1}, write u^{n+1} (12 B); the model,
    # density and strip weights are shared by the shots of a thread
    bytes_per_cell = 12.0
    step_s = loop_s / nt
    achieved = bytes_per_cell * B * nz * nx / step_s
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("k_rho_habc_fwd")
    out = {
        "metric": "Gcell-updates/s variable-density HABC forward",
        "value": round(cells * args.steps / el / 1e9, 3), "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded anomaly model, density 1000 + 1.2 kg/m^3 per cell of depth)",
        "config": {"workload": "variable-density acoustic HABC forward records",
                   "grid": [P["nz_i"], P["nx_i"]], "padded_grid": [nz, nx], "shots": B, "nt": nt},
        "roofline": {"bound": "hbm", "kernel": "k_rho_habc_fwd",
                     "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "avg_launch_us": round(step_s * 1e6, 2), "launches": args.steps * nt,
                     "bytes_per_cell_update": bytes_per_cell},
        "checksum": float(rec.double().abs().sum()),
    }
    # the gradient (vp and rho): forward with ring and snapshots, then the
    # boundary-saving adjoint (k_rho_adj, k_rho_inject, k_rho_grad per step)
    G = torch.randn_like(rec)
    vp = P["vp_init"].detach().clone().requires_grad_(True)
    rr = rho.detach().clone().requires_grad_(True)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    r = ar.propagate(vp, rr, b, P["w"], P["src"], P["rec"], B, P["dt"], P["h"])
    e1.record()
    r.backward(G)
    e2.record()
    torch.cuda.synchronize()
    f_s, a_s = e0.elapsed_time(e1) / 1e3, e1.elapsed_time(e2) / 1e3
    # adjoint per cell-update: lambda^{k+1}, lambda^{k+2} read and lambda^k
    # written (k_rho_adj), then lambda^k, u^{k-1}, u^k read and u^{k-2}
    # written (k_rho_grad): 28 B; model and gradient partials shared by shots
    adj_bytes = 28.0
    out["gradient"] = {
        "value": round(cells * 2 / (f_s + a_s) / 1e9, 3),
        "unit": "Gcell-updates/s fwd+adjoint (vp and rho gradients)",
        "forward_s": round(f_s, 3), "adjoint_s": round(a_s, 3),
        "adjoint_step_us": round(a_s / nt * 1e6, 1),
        "adjoint_roofline": {"bound": "hbm", "achieved": round(adj_bytes * B * nz * nx /
                                                               (a_s / nt) / 1e9, 1),
                             "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                             "frac": round(adj_bytes * B * nz * nx / (a_s / nt) / HBM_PEAK, 4),
                             "bytes_per_cell_update": adj_bytes},
        "grad_vp_abs_sum": float(vp.grad.double().abs().sum()),
        "grad_rho_abs_sum": float(rr.grad.double().abs().sum()),
    }
    if not args.no_cpu_baseline:
        from oracle import ref_torch
        steps = 40
        c_np, r_np, b_np = P["vp_true"].cpu().numpy(), rho.cpu().numpy(), b.cpu().numpy()
        t1 = time.perf_counter()
        ref_torch.rho_habc_forward(c_np, r_np, b_np, P["dt"], P["h"], P["w_np"][:steps],
                                   P["src_np"][:1], P["rec_np"][:1])
        cel = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(nz * nx * steps / cel / 1e9, 5),
                               "unit": "Gcell-updates/s", "cores": torch.get_num_threads(),
                               "cpu": cpu_model(), "kind": "port",
                               "sample": f"1 shot, {nz}x{nx} padded grid, {steps} of {nt} steps, "
                                         f"oracle/ref_torch.rho_habc_forward (FP32 torch on the "
                                         f"host), {cel:.1f} s"}
    print(json.dumps(out), flush=True)


def run_precond(args, dev):
    """Learned preconditioner: pairs from the generator, AdamW training with
    the hand-written kernels (or a saved network), then plain and
    preconditioned FWI on the 800x2400 grid with the 64-shot 8 Hz
    acquisition (workloads.precond_fwi: by default a true model of the
    generator's family at that size); value = wall-clock of the
    preconditioned run to reach the plain run's final misfit / the plain
    run's wall-clock.  Both runs take the same post-processed descent
    direction (source illumination, two Gaussian passes of sigma 4)."""
    import torch
    from seistorch_amd import fwi, learned, precond, workloads

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    p = precond.Params(dev, seed=0)
    net = precond.UNet(p)
    training = {}
    if args.load_net:
        p.load_state_dict(torch.load(args.load_net, map_location="cpu", weights_only=True))
        training = {"loaded": os.path.basename(args.load_net)}
        info = args.load_net + ".json"
        if os.path.exists(info):
            training.update(json.load(open(info)))
    else:
        prior = {}
        if args.init_net:  # continue training a saved network on further pairs
            p.load_state_dict(torch.load(args.init_net, map_location="cpu", weights_only=True))
            info = args.init_net + ".json"
            prior = json.load(open(info)) if os.path.exists(info) else {}
        t0 = time.perf_counter()
        trn = learned.make_pairs(args.pair_models, dev, first=args.pair_first, log=log)
        t_pairs = time.perf_counter() - t0
        # held-out pairs: the first models of a second seeded stream
        val = learned.make_pairs(args.val_models, dev, seed=HELD_OUT_SEED)
        opt = precond.AdamW(p, lr=1e-3)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        hist = learned.train(net, opt, trn, args.train_steps, log=log, log_every=500, flip=True,
                             seed=args.pair_first)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t1
        z, gs, nn, nb = learned.heldout_errors(net, val)
        seen = prior.get("pairs", 0) + len(trn)
        training = {"pairs": seen, "pairs_this_run": len(trn), "first_model": args.pair_first,
                    "held_out_pairs": len(val), "augmentation": "random x-mirror",
                    "pairs_s": round(prior.get("pairs_s", 0.0) + t_pairs, 1),
                    "train_steps": prior.get("train_steps", 0) + args.train_steps,
                    "train_s": round(prior.get("train_s", 0.0) + t_train, 1),
                    "loss_history": prior.get("loss_history", []) + hist,
                    "held_out_error_over_zero": {"best_multiple_of_gradient": round(gs / z, 4),
                                                 "network": round(nn / z, 4),
                                                 "best_multiple_of_network": round(nb / z, 4)}}
        if prior:
            training["runs"] = prior.get("runs", 1) + 1
        log(f"[bench] pairs {len(trn)} ({t_pairs:.1f} s), training {t_train:.1f} s, held-out "
            f"{training['held_out_error_over_zero']}")
        del val, trn
        if args.save_net:
            torch.save(p.state_dict(), args.save_net)
            json.dump(training, open(args.save_net + ".json", "w"))
    if args.fwi_iters == 0:  # training only
        print(json.dumps({"training": training}), flush=True)
        return
    W = workloads.precond_fwi(nt=args.nt, model=args.precond_model)
    from seistorch_amd.pml import generate_pml_coefficients_2d
    prob = fwi.Problem(torch.from_numpy(W["vp_true"]).to(dev),
                       torch.from_numpy(W["vp_init"]).to(dev),
                       generate_pml_coefficients_2d((W["nz"], W["nx"]), 50).to(dev),
                       torch.from_numpy(W["w"]).to(dev),
                       torch.from_numpy(W["src"].astype(np.int32)).to(dev),
                       torch.from_numpy(W["rec_flat"]).to(dev), W["B"], W["dt"], W["h"])
    prob.obs = prob.records(prob.vp_true)
    sm = {"counts": 2, "sigma": {"z": 4.0, "x": 4.0}, "radius": {"z": 8, "x": 8}}
    kw = dict(source_illumination=True, smooth=sm)
    _, hp = fwi.run_plain(prob, args.fwi_iters, log=log, **kw)
    _, hn = fwi.run_preconditioned(prob, net, args.fwi_iters, log=log, net_iters=1,
                                   scales=(1.0, 0.5, 0.25), **kw)
    target = hp.misfit[-1]
    t_hit = fwi.time_to(hn, target)
    ratio = (t_hit / hp.t[-1]) if t_hit is not None else None
    out = {
        "metric": METRIC, "value": None if ratio is None else round(ratio, 3),
        "unit": "preconditioned wall-clock to the plain run's final misfit / plain wall-clock",
        "n_gpus": 1, "steps": args.fwi_iters, "warmup": 0,
        "ms_per_step": round(hn.t[-1] / max(len(hn.t) - 1, 1) * 1e3, 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f32",
        "data": "synthetic (generator pairs for training; a seeded model for FWI)",
        "config": {"workload": "U-Net preconditioned FWI vs plain FWI, 800x2400, 64 shots, 8 Hz",
                   "fwi_model": args.precond_model, "fwi_iters": args.fwi_iters,
                   "nt": W["nt"], "descent_direction": "gradient / source illumination, "
                   "2 Gaussian passes sigma 4 cells"},
        "training": training,
        "plain": {"t": [round(x, 2) for x in hp.t], "misfit": hp.misfit, "rmse": hp.rmse},
        "preconditioned": {"t": [round(x, 2) for x in hn.t], "misfit": hn.misfit,
                           "rmse": hn.rmse, "kind": hn.kind},
        "misfit_at_plain_end": {"plain": target,
                                "preconditioned": fwi.misfit_at(hn, hp.t[-1])},
        "rmse_at_plain_end": {"plain": hp.rmse[-1], "preconditioned": hn.rmse[-1]},
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from seistorch_amd import _lib
    dev = torch.device("cuda:0")
    if args.workload == "generator":
        return run_generator(args, dev)
    if args.workload == "precond":
        return run_precond(args, dev)
    if args.workload == "filter":
        return run_filter(args, dev)
    if args.workload == "postprocess":
        return run_postprocess(args, dev)
    if args.workload == "rho":
        return run_rho(args, dev)
    elastic = args.workload == "elastic"
    if elastic:
        P = build_elastic_problem(dev, nt_override=args.nt)
        with torch.no_grad():
            P["obs"] = elastic_records(P, P["vp_true"], P["vs_true"], P["rho_true"])
        step_fn = elastic_gradient_step
    elif args.workload == "habc":
        from seistorch_amd.habc import generate_habc_coefficients_2d
        P = build_problem(dev, nt_override=args.nt)
        P["b"] = generate_habc_coefficients_2d((P["nz"], P["nx"]), 50).to(dev)
        with torch.no_grad():
            P["obs"] = habc_records(P, P["vp_true"])
        step_fn = habc_gradient_step
    else:
        P = build_problem(dev, nt_override=args.nt)
        with torch.no_grad():
            from seistorch_amd.equations2d import acoustic as ac
            P["obs"] = ac.propagate(P["vp_true"], P["d"], P["w"], P["src"], P["rec"], P["B"],
                                    P["dt"], P["h"])
        step_fn = gradient_step
    torch.cuda.synchronize()
    print(f"[bench] problem ready: {P['B']} shots, {P['nz']}x{P['nx']}, nt={P['nt']}",
          file=sys.stderr, flush=True)
    shape, tune = None, {}
    from seistorch_amd import autotune
    if args.workload == "acoustic":
        shape, tune = autotune.tune_acoustic(P["vp_init"], P["d"], P["w"], P["src"], P["rec"],
                                             P["B"], P["dt"], P["h"])
    elif args.workload == "elastic":
        shape, tune = autotune.tune_elastic(P["vp_init"], P["vs_init"], P["rho_init"], P["d"],
                                            P["w"], P["src"], P["rec"], P["B"], P["dt"], P["h"],
                                            src_type=("txx", "tzz"), rec_type=("vx", "vz"))
    elif args.workload == "habc":
        shape, tune = autotune.tune_habc(P["vp_init"], P["b"], P["w"], P["src"], P["rec"], P["B"],
                                         P["dt"], P["h"])
    if shape is not None:
        print(f"[bench] autotuned launch shape {shape}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        step_fn(P)
        torch.cuda.synchronize()
        print(f"[bench] warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    lib = _lib.load()
    lib.seis_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        loss, grad = step_fn(P)
        print(f"[bench] step {i + 1}/{args.steps} enqueued", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib.seis_profile_enable(0)
    kern = {}
    import ctypes
    for name, kid in KERNEL_IDS[args.workload]:
        ms = ctypes.c_double(0.0)
        cnt = ctypes.c_int64(0)
        _lib.call("seis_profile_collect", kid, ctypes.byref(ms), ctypes.byref(cnt))
        kern[name] = (ms.value, cnt.value)
    cells = P["B"] * P["nz"] * P["nx"] * P["nt"]
    value = cells * args.steps / el / 1e9
    dom = max(kern, key=lambda k: kern[k][0])
    tot_ms, n_launch = kern[dom]
    avg_s = tot_ms / max(n_launch, 1) / 1e3
    alg_bytes = BYTES_PER_CELL[dom] * P["B"] * P["nz"] * P["nx"]
    achieved = alg_bytes / avg_s if avg_s > 0 else 0.0
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(dom)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded anomaly model, records from the true model)",
        "config": {"workload": {"elastic": "elastic vp/vs/rho gradient, PML 50, fwd + "
                                           "boundary-saving adjoint",
                                "habc": "acoustic FWI gradient, HABC 50, fwd + boundary-saving "
                                        "adjoint",
                                "acoustic": "acoustic FWI gradient, PML 50, fwd + "
                                            "boundary-saving adjoint"}[args.workload],
                   "grid": [P["nz_i"], P["nx_i"]], "padded_grid": [P["nz"], P["nx"]],
                   "shots": P["B"], "nt": P["nt"], "dt": P["dt"], "dx": P["h"],
                   "receivers_per_shot": P["nx_i"], "f_peak_hz": P["fm"]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "avg_launch_us": round(avg_s * 1e6, 2), "launches": n_launch,
                     "bytes_per_cell_update": BYTES_PER_CELL[dom]},
        "kernels_ms_total": {k: round(v[0], 2) for k, v in kern.items()},
        "launch_shape": None if shape is None else _launch_shape_json(args.workload, shape, tune),
        "loss": float(loss),
    }
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(dev)
    if not args.no_cpu_baseline and args.workload == "acoustic":
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(P, args.cpu_steps)
    elif not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_ad(P, args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
