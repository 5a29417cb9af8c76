"""Parity of the HIP gradient post-processing (seis_gaussian_filter,
seis_smooth_gradient) with the FP64 oracle and with the reference's own
steps restated in FP32 torch.

The kernels accumulate the taps in FP32 in tap order; the reference runs
torch's conv2d (its own summation order), so both are compared with a
tolerance: the FP32 rounding of a normalised ks-tap sum is below
ks * 2^-24 of the field's peak per pass."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import process as opp

pytestmark = pytest.mark.gpu


def _fp32_reference_steps(x, sigma, radius, axis):
    """signal.py:262-301 as written there, FP32 on the CPU."""
    from seistorch_amd.process import gaussian_kernel
    w = torch.from_numpy(gaussian_kernel(sigma, radius))
    ks = w.numel()
    pad = ks // 2
    xp = torch.from_numpy(np.pad(x, ((pad, pad), (pad, pad)), mode="reflect")).float()
    k, p = (w.view(1, 1, -1, 1), (pad, 0)) if axis == 0 else (w.view(1, 1, 1, -1), (0, pad))
    r = F.conv2d(xp[None, None], k, padding=p)[0, 0]
    return (r[pad:-pad, pad:-pad] if pad > 0 else r).numpy()


@pytest.mark.parametrize("shape", [(37, 53), (5, 300), (300, 6)])
@pytest.mark.parametrize("radius,sigma", [(0, 1.0), (1, 1.0), (2, 0.7), (3, 1.5), (4, 2.5),
                                          (9, 4.0)])
def test_gaussian_filter_matches_oracle(shape, radius, sigma):
    from seistorch_amd import process as pp
    x = np.random.default_rng(radius + shape[0]).standard_normal(shape).astype(np.float32)
    w32 = pp.gaussian_kernel(sigma, radius)
    for axis in (0, 1):
        if w32.size // 2 >= shape[axis]:
            with pytest.raises(RuntimeError, match="smaller than the axis"):
                pp.gaussian_filter(torch.from_numpy(x).cuda(), sigma, radius, axis)
            continue
        got = pp.gaussian_filter(torch.from_numpy(x).cuda(), sigma, radius, axis).cpu().numpy()
        ref = opp.gaussian_filter_taps(x, w32, axis)
        ref32 = _fp32_reference_steps(x, sigma, radius, axis)
        assert got.shape == ref.shape == ref32.shape
        peak = np.abs(x).max()
        e64, e32 = np.abs(got - ref).max() / peak, np.abs(got - ref32).max() / peak
        print(f"gaussian_filter shape={shape} r={radius} sigma={sigma} axis={axis} "
              f"out={got.shape}: vs FP64 oracle {e64:.2e}, vs FP32 reference steps {e32:.2e}")
        assert e64 <= w32.size * 2.0 ** -24 and e32 <= 2 * w32.size * 2.0 ** -24


def _bound(g, counts, kz, kx, illum=None):
    """Worst-case FP32 error of the pass sequence: each tap sum rounds at most
    ks times by 2^-24 of the running magnitude (<= max|g|: the taps are
    positive and sum to 1), plus the multiply and the division; the division
    scales everything by up to 1/min(illum)."""
    amp = 1.0 / float(illum.min()) if illum is not None else 1.0
    return 2 * (counts * (kz + kx) + 2) * 2.0 ** -24 * float(np.abs(g).max()) * amp


def _problem(nz, nx, seed):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nz, nx)).astype(np.float32)
    mask = np.ones((nz, nx), np.float32)
    mask[: nz // 3] = 0.0           # water layer above the seabed
    mask[nz // 3, ::3] = 0.5        # a ragged seabed row
    illum = (rng.random((nz, nx)) + 0.1).astype(np.float32)
    return g, mask, illum


@pytest.mark.parametrize("counts,rz,rx", [(1, 2, 2), (3, 4, 2), (2, 0, 6), (0, 2, 2)])
@pytest.mark.parametrize("use_mask,use_illum", [(False, False), (True, False), (True, True)])
def test_smooth_gradient_matches_oracle(counts, rz, rx, use_mask, use_illum):
    from seistorch_amd import process as pp
    g, mask, illum = _problem(61, 97, counts + rz)
    sm = {"counts": counts, "sigma": {"z": 1.5, "x": 2.0}, "radius": {"z": rz, "x": rx}}
    d = lambda a: torch.from_numpy(a).cuda()
    m = d(mask) if use_mask else None
    got = pp.smooth_gradient(d(g), sm, pre_mask=m, post_mask=m,
                             illum=d(illum) if use_illum else None).cpu().numpy()
    taps = (pp.gaussian_kernel(1.5, rz), pp.gaussian_kernel(2.0, rx))
    ref = opp.postprocess(g, sm, pre_mask=mask if use_mask else None,
                          post_mask=mask if use_mask else None,
                          illum=illum if use_illum else None, taps=taps)
    err = np.abs(got - ref).max()
    bound = _bound(g, counts, taps[0].size, taps[1].size, illum if use_illum else None)
    print(f"smooth_gradient counts={counts} r=({rz},{rx}) mask={use_mask} illum={use_illum}: "
          f"max abs err {err:.2e} (bound {bound:.2e}, output peak {np.abs(ref).max():.2e})")
    assert err <= bound
    if use_mask:
        assert np.all(got[: 61 // 3] == 0.0)


def test_postprocess_methods_equal_the_fused_call():
    """PostProcess.smooth_gradient -> cut_gradient -> precondition (three
    calls, the reference's order) equals apply() (one sequence) bit for bit,
    and both match the oracle."""
    from seistorch_amd import process as pp
    g, seabed, illum = _problem(50, 80, 3)
    pad = 5
    seabed_in = seabed[pad:-pad, pad:-pad]
    cfg = {"geom": {"boundary": {"width": pad}, "multiple": False, "seabed": seabed_in},
           "training": {"smooth": {"counts": 2, "sigma": {"z": 1.0, "x": 1.5},
                                   "radius": {"z": 2, "x": 4}}}}

    class Cmd:
        grad_cut = True

    def model():
        p = torch.nn.Parameter(torch.zeros(50, 80, device="cuda"))
        p.grad = torch.from_numpy(g).cuda()
        return pp.Params([p], precondition=torch.from_numpy(illum).cuda())

    m1 = model()
    post = pp.PostProcess(m1, cfg, Cmd())
    post.smooth_gradient()
    post.cut_gradient()
    post.precondition()
    m2 = model()
    pp.PostProcess(m2, cfg, Cmd()).apply(grad_smooth=True, grad_cut=True,
                                         source_illumination=True)
    a, b = next(m1.parameters()).grad, next(m2.parameters()).grad
    assert torch.equal(a, b)
    mask = opp.seabed_mask(seabed_in, pad, multiple=False)
    sm = cfg["training"]["smooth"]
    ref = opp.postprocess(g, sm, pre_mask=mask, post_mask=mask, illum=illum,
                          taps=(pp.gaussian_kernel(1.0, 2), pp.gaussian_kernel(1.5, 4)))
    err = np.abs(a.cpu().numpy() - ref).max()
    bound = _bound(g, 2, 5, 9, illum)
    print(f"PostProcess methods == apply(): True; vs oracle {err:.2e} (bound {bound:.2e})")
    assert err <= bound


def test_odd_radius_is_rejected_by_the_gradient_path():
    from seistorch_amd import process as pp
    g = torch.zeros(20, 20, device="cuda")
    sm = {"counts": 1, "sigma": {"z": 1.0, "x": 1.0}, "radius": {"z": 1, "x": 2}}
    with pytest.raises(RuntimeError, match="must be odd"):
        pp.smooth_gradient(g, sm)


def test_full_size_properties():
    """C2-sized gradient (900 x 2500): a constant stays constant through
    five passes per axis (taps sum to 1 in FP32), and the cut region is
    exactly zero."""
    from seistorch_amd import process as pp
    sm = {"counts": 5, "sigma": {"z": 3.0, "x": 3.0}, "radius": {"z": 10, "x": 10}}
    c = torch.full((900, 2500), 2.5, device="cuda")
    out = pp.smooth_gradient(c, sm)
    dev = (out - 2.5).abs().max().item()
    mask = torch.ones(900, 2500, device="cuda")
    mask[:100] = 0
    cut = pp.smooth_gradient(torch.randn(900, 2500, device="cuda"), sm, pre_mask=mask,
                             post_mask=mask)
    print(f"900x2500 constant after 5+5 passes: max deviation {dev:.2e}")
    assert dev <= 10 * 21 * 2.0 ** -24 * 2.5
    assert bool((cut[:100] == 0).all())


@pytest.mark.parametrize("shape,rz,rx", [((61, 97), 2, 4), ((900, 2500), 10, 10), ((40, 300), 16, 0)])
def test_one_count_equals_two_single_axis_passes(shape, rz, rx):
    """One smoothing count of seis_smooth_gradient equals the z pass then the
    x pass of the single-axis entry point bit for bit (same operands, same
    order, same FP32 intermediate), at small and full sizes."""
    from seistorch_amd import process as pp
    g = torch.randn(shape, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    sm = {"counts": 1, "sigma": {"z": 2.0, "x": 3.0}, "radius": {"z": rz, "x": rx}}
    fused = pp.smooth_gradient(g, sm)
    two = pp.gaussian_filter(pp.gaussian_filter(g, 2.0, rz, 0), 3.0, rx, 1)
    print(f"one count {shape} r=({rz},{rx}): equal to two passes {torch.equal(fused, two)}")
    assert torch.equal(fused, two)


@pytest.mark.parametrize("multiple", [False, True])
def test_fwi_postprocess_on_the_padded_gradient(multiple):
    """fwi.postprocess_padded against the reference's own steps on the padded
    gradient (seistorch_dist.py:270-280, process.py:26-117): seabed mask of
    the unpadded grid zero-padded (no top pad with a free surface), g * mask,
    counts z and x passes (numpy reflect pad -> conv2d -> crop, FP32), g *
    mask, g / illumination, then the interior crop.  Cells near the interior
    edges see the strips' values through the smoothing."""
    from seistorch_amd import fwi
    P = 50
    nz_i, nx_i = 40, 70
    nz, nx = nz_i + (P if multiple else 2 * P), nx_i + 2 * P
    rng = np.random.default_rng(3)
    g = np.zeros((nz, nx), np.float32)
    top = 0 if multiple else P
    g[top:top + nz_i, P:P + nx_i] = rng.standard_normal((nz_i, nx_i)).astype(np.float32)
    seabed = np.ones((nz_i, nx_i), np.float32)
    seabed[:6] = 0.0
    illum = (rng.random((nz, nx)) + 0.5).astype(np.float32)
    smooth = {"counts": 2, "sigma": {"z": 2.0, "x": 3.0}, "radius": {"z": 4, "x": 6}}
    got = fwi.postprocess_padded(torch.from_numpy(g).cuda(), smooth,
                                 torch.from_numpy(seabed).cuda(), torch.from_numpy(illum).cuda(),
                                 multiple=multiple).cpu().numpy()
    mask = np.pad(seabed, ((top, P), (P, P)))
    ref = g * mask
    for _ in range(smooth["counts"]):
        ref = _fp32_reference_steps(ref, 2.0, 4, 0)
        ref = _fp32_reference_steps(ref, 3.0, 6, 1)
    ref = ref * mask / illum
    ref = ref[top:nz - P, P:nx - P]
    assert got.shape == ref.shape
    e = np.abs(got - ref).max() / np.abs(ref).max()
    edge = np.abs(got[:, :8] - ref[:, :8]).max() / np.abs(ref).max()
    print(f"fwi post-processing on the padded gradient, multiple={multiple}: max {e:.2e} of the "
          f"peak (first 8 interior columns {edge:.2e})")
    assert e < 1e-5
