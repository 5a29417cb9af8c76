"""The seeded uneven generator on the CPU: instances drawn from any point of
the stream equal the same instances of the whole stream (the learned
preconditioner's training runs draw successive slices of one stream), and
the drawn parameters stay in the configured ranges."""
import numpy as np

from seistorch_amd import generator


def test_slices_of_the_stream_equal_the_whole_stream():
    spec = generator.GenSpec(n_models=5, nz=(64, 96), nx=(64, 96))
    whole = generator.make_instances(seed=7, spec=spec)
    part = generator.make_instances(seed=7, spec=generator.GenSpec(
        **{**spec.__dict__, "n_models": 2}), first=3)
    for a, b in zip(whole[3:], part):
        assert a["seed"] == b["seed"] and a["B"] == b["B"] and a["nt"] == b["nt"]
        assert np.array_equal(a["vp_true"], b["vp_true"])
        assert np.array_equal(a["vp_start"], b["vp_start"])
        assert np.array_equal(a["rec_flat"], b["rec_flat"])


def test_instances_in_range():
    spec = generator.GenSpec(n_models=6, nz=(64, 128), nx=(64, 160))
    for it in generator.make_instances(seed=3, spec=spec):
        assert spec.nz[0] <= it["nz_i"] <= spec.nz[1]
        assert spec.nx[0] <= it["nx_i"] <= spec.nx[1] and it["nx_i"] % spec.nx_step == 0
        assert it["vp_true"].min() >= spec.vmin and it["vp_true"].max() <= spec.vmax
        assert spec.shots[0] <= it["B"] <= spec.shots[1]
        assert it["vp_true"].shape == (it["nz_i"] + 100, it["nx_i"] + 100)
