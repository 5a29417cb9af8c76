"""FP32 op-for-op torch restatement of the reference acoustic PML path.

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py's
cpu_baseline leg; never by the product.

It uses the same torch ops in the same order as the reference so that its
FP32 rounding behaves like the reference's CPU path:
* Laplacian  conv2d(u[:, None], h**-2 * K, padding=1)       acoustic.py:60-71
* step       (dt**-2 + b*dt**-1).pow(-1) * (2/dt**2*h1
               - (dt**-2 - b*dt**-1)*h2 + c.pow(2)*lap)       acoustic.py:81-84
* source     Y = Y.clone(); Y += smask * w[n]                 source.py:49-52
* receivers  u[bidx, z, x]                                    probe.py:42-44
* loop       step -> source -> record                        rnn.py:178-212
* misfit     sum of MSELoss(reduction='sum') over shots       loss.py:417-421

``gradient_ad`` backpropagates through the whole loop (the reference's
``boundary_saving: false`` path, cell.py:74-75).  ``gradient_bs`` trades
memory for recompute the way the reference's boundary-saving path does
(checkpoint_new.py:114-217): the forward keeps only the 1-cell ring at
index 49/-50 per step, the backward rebuilds the interior in reverse time and
differentiates one recomputed step at a time.  It follows the exact
time pairing (not the reference's two-step shift, which is a defect of that
path) so its interior gradient equals ``gradient_ad``'s.
"""
import torch
import torch.nn.functional as F

_K = torch.tensor([[0.0, 1.0, 0.0], [1.0, -4.0, 1.0], [0.0, 1.0, 0.0]])


def laplacian(y, h):
    op = (h ** (-2) * _K.to(device=y.device, dtype=y.dtype)).view(1, 1, 3, 3)
    return F.conv2d(y.unsqueeze(1), op, padding=1).squeeze(1)


def time_step(c, h1, h2, dt, h, b):
    y = torch.mul((dt ** -2 + b * dt ** -1).pow(-1),
                  (2 / dt ** 2 * h1 - torch.mul((dt ** -2 - b * dt ** -1), h2)
                   + torch.mul(c.pow(2), laplacian(h1, h))))
    return y, h1


def _setup(c, d, dt, h, wavelet, src, rec, dtype, device=None):
    """Tensors of one problem; device=None keeps them on the CPU (a HIP
    device runs the same torch ops there, e.g. for full-size problems)."""
    c = torch.as_tensor(c).to(device=device, dtype=dtype)
    d = torch.as_tensor(d).to(device=device, dtype=dtype)
    dt_t = torch.tensor(dt, dtype=dtype, device=device)
    h_t = torch.tensor(h, dtype=dtype, device=device)
    w = torch.as_tensor(wavelet).to(device=device, dtype=dtype)
    src = torch.as_tensor(src, dtype=torch.int64).reshape(-1, 2).to(device)
    recs = [torch.as_tensor(r, dtype=torch.int64).reshape(-1, 2).to(device) for r in rec]
    B = src.shape[0]
    nz, nx = c.shape
    smask = torch.zeros((B, nz, nx), dtype=dtype, device=device)
    smask[torch.arange(B, device=device), src[:, 0], src[:, 1]] = 1.0
    bidx = torch.cat([torch.full((r.shape[0],), b, dtype=torch.int64, device=device)
                      for b, r in enumerate(recs)])
    rz = torch.cat([r[:, 0] for r in recs])
    rx = torch.cat([r[:, 1] for r in recs])
    counts = [r.shape[0] for r in recs]
    return c, d, dt_t, h_t, w, smask, (bidx, rz, rx), counts


def forward(c, d, dt, h, wavelet, src, rec, dtype=torch.float32, c_param=None, device=None):
    """Records as a list of (nt, nrec_b, 1) tensors (rnn.py:207-212)."""
    c0, d, dt_t, h_t, w, smask, (bidx, rz, rx), counts = _setup(c, d, dt, h, wavelet, src, rec,
                                                                dtype, device)
    c = c0 if c_param is None else c_param
    B = smask.shape[0]
    h1 = torch.zeros_like(smask)
    h2 = torch.zeros_like(smask)
    recs = []
    for n in range(w.shape[0]):
        h1, h2 = time_step(c, h1, h2, dt_t, h_t, d)
        y = h1.clone()
        y += smask * w[n]
        h1 = y
        recs.append(h1[bidx, rz, rx])
    stacked = torch.stack(recs, 0).unsqueeze(2)
    return list(torch.split(stacked, counts, dim=1))


def l2(syn, obs):
    loss = 0.0
    for s, o in zip(syn, obs):
        loss = loss + torch.nn.MSELoss(reduction='sum')(s, torch.as_tensor(o).to(s.dtype))
    return loss


def gradient_ad(c, d, dt, h, wavelet, src, rec, obs, dtype=torch.float32, rec_cotangent=None):
    """Autograd through the whole loop: returns (J, dJ/dc, records).

    With ``rec_cotangent`` (nt, sum nrec) the records are backpropagated with
    that cotangent instead of the L2 residual (J is then returned as 0).
    """
    cp = torch.as_tensor(c).to(dtype).clone().requires_grad_(True)
    syn = forward(c, d, dt, h, wavelet, src, rec, dtype=dtype, c_param=cp)
    if rec_cotangent is None:
        J = l2(syn, obs)
        J.backward()
    else:
        J = torch.zeros(())
        torch.cat(syn, 1)[..., 0].backward(torch.as_tensor(rec_cotangent).to(dtype))
    return J.detach(), cp.grad.detach(), [s.detach() for s in syn]


def _ring_idx(nz, nx, npml=49, multiple=False):
    """Ring cells (z, x) of seistorch/equations2d/utils.py:50-90 without the
    corners; with a free surface the top row is not part of it and the
    columns reach row 0 (utils.py:71-78 leaves the top alone)."""
    z0, z1, x0, x1 = npml, nz - 1 - npml, npml, nx - 1 - npml
    xs = torch.arange(x0 + 1, x1)
    zs = torch.arange(0 if multiple else z0 + 1, z1)
    top_z = [] if multiple else [torch.full_like(xs, z0)]
    top_x = [] if multiple else [xs]
    z = torch.cat(top_z + [torch.full_like(xs, z1), zs, zs])
    x = torch.cat(top_x + [xs, torch.full_like(zs, x0), torch.full_like(zs, x1)])
    return z, x


def gradient_bs(c, d, dt, h, wavelet, src, rec, obs, dtype=torch.float32, npml=49,
                multiple=False, rec_cotangent=None, habc=False, device=None):
    """Boundary-saving gradient by per-step recompute (CPU baseline shape).

    Returns (J, dJ/dc on the interior [npml+1, -npml-1) -- rows from 0 with a
    free surface -- and records).  With ``rec_cotangent`` (nt, sum nrec) the
    records are backpropagated with it instead of the L2 residual (J = 0).
    habc=True: ``d`` is the HABC weight field and every step is
    ``habc_step`` (acoustic_habc.py:206-221); the rebuilt interior is the
    same undamped region.
    """
    c, d, dt_t, h_t, w, smask, (bidx, rz, rx), counts = _setup(c, d, dt, h, wavelet, src, rec,
                                                               dtype, device)
    nz, nx = c.shape
    nt = w.shape[0]
    rzi, rxi = (i.to(c.device) for i in _ring_idx(nz, nx, npml, multiple))

    def step(cc, h1, h2):
        if habc:
            return habc_step(cc, h1, h2, dt_t, h_t, d, multiple=multiple)
        return time_step(cc, h1, h2, dt_t, h_t, d)

    rings = []
    h1 = torch.zeros_like(smask)
    h2 = torch.zeros_like(smask)
    recs = []
    with torch.no_grad():
        for n in range(nt):
            h1, h2 = step(c, h1, h2)
            h1 = h1 + smask * w[n]
            recs.append(h1[bidx, rz, rx])
            rings.append(h1[:, rzi, rxi].clone())
    stacked = torch.stack(recs, 0).unsqueeze(2)
    syn = list(torch.split(stacked, counts, dim=1))
    if rec_cotangent is None:
        J = l2(syn, obs)
        res = 2.0 * (stacked - torch.cat([torch.as_tensor(o).to(dtype) for o in obs], 1))[..., 0]
    else:
        J = torch.zeros(())
        res = torch.as_tensor(rec_cotangent).to(device=c.device, dtype=dtype)
    grad = torch.zeros_like(c)
    # state: u_cur = u^{n+1}, u_prev = u^n ; adjoint grads wrt them
    u_next, u_cur = h1, h2
    g_next = torch.zeros_like(smask)
    g_cur = torch.zeros_like(smask)
    lo_z = 0 if multiple else npml + 1
    lo, hi_z, hi_x = npml + 1, nz - npml - 1, nx - npml - 1
    for n in range(nt - 1, -1, -1):
        # adjoint of the record of u^{n+1}
        g_next = g_next.index_put((bidx, rz, rx), res[n], accumulate=True)
        # rebuild u^{n-1} on the interior from u^{n+1}, u^n (source removed)
        with torch.no_grad():
            u_src = u_next - smask * w[n]
            prev = torch.zeros_like(u_cur)
            if n > 1:  # u^0 = u^-1 = 0; rings[m] holds the ring of u^{m+1}
                lap = laplacian(u_cur, h_t)
                prev = 2 * u_cur - u_src + (c.pow(2) * dt_t ** 2) * lap
                keep = torch.zeros_like(prev)
                keep[:, lo_z:hi_z, lo:hi_x] = prev[:, lo_z:hi_z, lo:hi_x]
                keep[:, rzi, rxi] = rings[n - 2]
                prev = keep
        cp = c.clone().requires_grad_(True)
        ucur = u_cur.clone().requires_grad_(True)
        uprev = prev.clone().requires_grad_(True)
        y, _ = step(cp, ucur, uprev)
        y.backward(g_next)
        grad += cp.grad
        g_next, g_cur = g_cur + ucur.grad, uprev.grad
        u_next, u_cur = u_cur, prev
    mask = torch.zeros_like(grad)
    mask[lo_z:hi_z, lo:hi_x] = 1.0
    return J.detach(), grad * mask, syn


# ------------------------------------------------------------------ elastic

def diff_using_roll(x, dim=-1, forward=True, padding_value=0):
    """seistorch/equations2d/utils.py:3-48: forward -> x - roll(x, 1) with the
    first index zeroed; backward -> roll(x, -1) - x with the last zeroed."""
    if forward:
        diff = x - torch.roll(x, shifts=1, dims=dim)
        if dim == 1:
            diff[:, 0] = padding_value
        elif dim == 2:
            diff[..., 0] = padding_value
    else:
        diff = torch.roll(x, shifts=-1, dims=dim) - x
        if dim == 1:
            diff[:, -1] = padding_value
        elif dim == 2:
            diff[..., -1] = padding_value
    return diff


def elastic_step(vp, vs, rho, vx, vz, txx, tzz, txz, dt, h, d):
    """seistorch/equations2d/elastic.py:7-37, same expressions and order."""
    lame_lambda = rho * (vp.pow(2) - 2 * vs.pow(2))
    lame_mu = rho * (vs.pow(2))
    c = 0.5 * dt * d
    vx_x = diff_using_roll(vx, 2)
    vz_z = diff_using_roll(vz, 1, False)
    vx_z = diff_using_roll(vx, 1)
    vz_x = diff_using_roll(vz, 2, False)
    y_txx = (1 + c) ** -1 * (dt * h.pow(-1) * ((lame_lambda + 2 * lame_mu) * vx_x + lame_lambda * vz_z) + (1 - c) * txx)
    y_tzz = (1 + c) ** -1 * (dt * h.pow(-1) * ((lame_lambda + 2 * lame_mu) * vz_z + lame_lambda * vx_x) + (1 - c) * tzz)
    y_txz = (1 + c) ** -1 * (dt * lame_mu * h.pow(-1) * (vz_x + vx_z) + (1 - c) * txz)
    txx_x = diff_using_roll(y_txx, 2, False)
    txz_z = diff_using_roll(y_txz, 1, False)
    tzz_z = diff_using_roll(y_tzz, 1)
    txz_x = diff_using_roll(y_txz, 2)
    y_vx = (1 + c) ** -1 * (dt * rho.pow(-1) * h.pow(-1) * (txx_x + txz_z) + (1 - c) * vx)
    y_vz = (1 + c) ** -1 * (dt * rho.pow(-1) * h.pow(-1) * (txz_x + tzz_z) + (1 - c) * vz)
    return y_vx, y_vz, y_txx, y_tzz, y_txz


ELASTIC_FIELDS = ("vx", "vz", "txx", "tzz", "txz")


def elastic_forward(vp, vs, rho, d, dt, h, wavelet, src, rec, src_type=("txx", "tzz"),
                    rec_type=("vx", "vz"), dtype=torch.float32, params=None, device=None):
    """Records as a list of (nt, nrec_b, nch) tensors (rnn.py:176-212, offset 0)."""
    vp0, d, dt_t, h_t, w, smask, (bidx, rz, rx), counts = _setup(vp, d, dt, h, wavelet, src, rec,
                                                                 dtype, device)
    if params is None:
        params = (vp0, torch.as_tensor(vs).to(device=device, dtype=dtype),
                  torch.as_tensor(rho).to(device=device, dtype=dtype))
    f = {n: torch.zeros_like(smask) for n in ELASTIC_FIELDS}
    recs = {k: [] for k in rec_type}
    for n in range(w.shape[0]):
        new = elastic_step(*params, *(f[x] for x in ELASTIC_FIELDS), dt_t, h_t, d)
        f = dict(zip(ELASTIC_FIELDS, new))
        for st in src_type:
            y = f[st]
            y += smask * w[n]  # first-order equations add in place (source.py:49-52)
        for k in rec_type:
            recs[k].append(f[k][bidx, rz, rx])
    stacked = torch.stack([torch.stack(recs[k], 0) for k in rec_type], 2)
    return list(torch.split(stacked, counts, dim=1))


def elastic_gradient_ad(vp, vs, rho, d, dt, h, wavelet, src, rec, obs, src_type=("txx", "tzz"),
                        rec_type=("vx", "vz"), dtype=torch.float32, rec_cotangent=None):
    """Autograd through the whole loop: (J, (g_vp, g_vs, g_rho), records)."""
    ps = [torch.as_tensor(x).to(dtype).clone().requires_grad_(True) for x in (vp, vs, rho)]
    syn = elastic_forward(vp, vs, rho, d, dt, h, wavelet, src, rec, src_type, rec_type, dtype,
                          params=ps)
    if rec_cotangent is None:
        J = l2(syn, obs)
        J.backward()
    else:
        J = torch.zeros(())
        torch.cat(syn, 1).backward(torch.as_tensor(rec_cotangent).to(dtype))
    return J.detach(), tuple(p.grad.detach() for p in ps), [s.detach() for s in syn]


# --------------------------------------------------------------------- HABC

def habc_masks(nz, nx, w=50, multiple=False):
    """Trapezoid strip masks (seistorch/habc.py:4-40): top (w, nx), bottom (w, nx),
    left (nz, w), right (nz, w); top is None with a free surface."""
    i = torch.arange(w)[:, None]
    j = torch.arange(nx)[None, :]
    top = (j >= i) & (j <= nx - 1 - i)
    bottom = torch.flip(top, [0])
    r = torch.arange(nz)[:, None]
    k = torch.arange(w)[None, :]
    left = (k <= r) & (k <= nz - 1 - r)
    right = torch.flip(left, [1])
    if multiple:
        top = None
        left = left.clone(); right = right.clone()
        left[:w] = True
        right[:w] = True
    return top, bottom, left, right


def _one_way_strips(y, h1, h2, c, b, dt, h, w):
    """Strips stacked (2, B, w+1, n) with row 0 at the outer edge, one-way
    extrapolation along the row axis with torch.roll (acoustic_habc.py:79-101)."""
    lam = 2 * c * dt / h
    mu = c ** 2 * dt ** 2 / (h ** 2)
    t1 = (2 - lam - mu) * h1
    t2 = (lam + 2 * mu) * torch.roll(h1, shifts=-1, dims=2)
    t3 = -mu * torch.roll(h1, shifts=-2, dims=2)
    t4 = (lam - 1) * h2
    t5 = -lam * torch.roll(h2, shifts=-1, dims=2)
    u_one = t1 + t2 + t3 + t4 + t5
    return (u_one * b + (1 - b) * y)[:, :, :w, :]


def habc_step(c, h1, h2, dt, h, b, w=50, multiple=False):
    """FP32 restatement of acoustic_habc._time_step (:206-221) with habc() (:147-204)."""
    y = torch.mul((dt ** -2).pow(-1), (2 / dt ** 2 * h1 - torch.mul((dt ** -2), h2)
                                       + torch.mul(c.pow(2), laplacian(h1, h))))
    return habc_blend(y, h1, h2, c, dt, h, b, w, multiple), h1


def habc_blend(y, h1, h2, c, dt, h, b, w=50, multiple=False):
    """habc() of acoustic_habc.py:147-204: the strips of y replaced by the
    weighted one-way extrapolation of (h1, h2), corners averaged."""
    Bn, nz, nx = y.shape
    fields = [y, h1, h2, c.unsqueeze(0).expand(Bn, nz, nx), b.unsqueeze(0).expand(Bn, nz, nx)]
    # top / bottom strips: rows from the edge inwards
    tb = [torch.stack([f[:, :w + 1, :], torch.flip(f, [1])[:, :w + 1, :]]) for f in fields]
    # left / right strips: columns from the edge inwards, as rows
    lr = [torch.stack([f.transpose(1, 2)[:, :w + 1, :],
                       torch.flip(f.transpose(1, 2), [1])[:, :w + 1, :]]) for f in fields]
    tbv = _one_way_strips(*tb, dt, h, w)
    lrv = _one_way_strips(*lr, dt, h, w)
    y_top = tbv[0]                                  # (B, w, nx), rows 0..w-1
    y_bottom = torch.flip(tbv[1], [1])              # rows nz-w..nz-1
    y_left = lrv[0].transpose(1, 2)                 # (B, nz, w), cols 0..w-1
    y_right = torch.flip(lrv[1], [1]).transpose(1, 2)  # cols nx-w..nx-1
    top, bottom, left, right = habc_masks(nz, nx, w, multiple)
    y = y.clone()
    if top is not None:
        y[:, :w, :][:, top] = y_top[:, top]
    y[:, -w:, :][:, bottom] = y_bottom[:, bottom]
    y[:, :, :w][:, left] = y_left[:, left]
    y[:, :, -w:][:, right] = y_right[:, right]
    d = torch.arange(w)
    if not multiple:
        y[:, d, d] = 0.5 * y_top[:, d, d] + 0.5 * y_left[:, d, d]
        y[:, d, nx - w + d] = 0.5 * y_top[:, d, nx - w + d] + 0.5 * y_right[:, d, d]
    y[:, nz - w + d, d] = 0.5 * y_bottom[:, d, d] + 0.5 * y_left[:, nz - w + d, d]
    y[:, nz - w + d, nx - w + d] = 0.5 * y_bottom[:, d, nx - w + d] + 0.5 * y_right[:, nz - w + d, d]
    return y


def _shift(p, dz, dx):
    """p at (z + dz, x + dx), zero outside the grid (the zero padding of the
    reference's conv2d with padding 1)."""
    nz, nx = p.shape[-2:]
    out = torch.zeros_like(p)
    zs, zd = (slice(dz, nz), slice(0, nz - dz)) if dz >= 0 else (slice(0, nz + dz), slice(-dz, nz))
    xs, xd = (slice(dx, nx), slice(0, nx - dx)) if dx >= 0 else (slice(0, nx + dx), slice(-dx, nx))
    out[..., zd, xd] = p[..., zs, xs]
    return out


def rho_operator(rho, p):
    """The bracket of acoustic_rho_habc.py:37-53: sum over x and z of
    b_pos * (p[+1] - p) - b_neg * (p - p[-1]), with b = 1/rho averaged over
    the two cells of each face (Geiger and Daley 2003, eq. 9), zero outside
    the grid for both p and b."""
    bu = rho.pow(-1)
    pxp = _shift(p, 0, 1) - p
    pxn = p - _shift(p, 0, -1)
    pzp = _shift(p, 1, 0) - p
    pzn = p - _shift(p, -1, 0)
    bxp = (bu + _shift(bu, 0, 1)) / 2.0
    bxn = (_shift(bu, 0, -1) + bu) / 2.0
    bzp = (bu + _shift(bu, 1, 0)) / 2.0
    bzn = (_shift(bu, -1, 0) + bu) / 2.0
    return bxp * pxp - bxn * pxn + bzp * pzp - bzn * pzn


def rho_habc_step(c, rho, h1, h2, dt, h, b, w=50, multiple=False):
    """acoustic_rho_habc._time_step (:31-57): p' = 2 p1 - p2 +
    dt^2 h^-2 rho c^2 * bracket (dt, h Python floats: the constant is formed
    in double, as the reference's configuration scalars are), then habc()."""
    C = float(dt) ** 2 * float(h) ** (-2)
    y = 2 * h1 - h2 + C * rho * c ** 2 * rho_operator(rho, h1)
    # the strips take dt and h as tensors of the field's precision, as
    # habc_step does (the one-way coefficients square dt in that precision)
    dt_t = torch.tensor(float(dt), dtype=c.dtype)
    h_t = torch.tensor(float(h), dtype=c.dtype)
    return habc_blend(y, h1, h2, c, dt_t, h_t, b, w, multiple), h1


def rho_habc_forward(c, rho, b, dt, h, wavelet, src, rec, dtype=torch.float32, multiple=False,
                     params=None):
    """Records of the variable-density HABC loop (rnn.py:178-212), list of
    (nt, nrec_b, 1).  params: optional (c, rho) tensors to differentiate."""
    c, b, _, _, w, smask, (bidx, rz, rx), counts = _setup(c, b, dt, h, wavelet, src, rec, dtype)
    rho = torch.as_tensor(rho).to(dtype)
    if params is not None:
        c, rho = params
    h1 = torch.zeros_like(smask)
    h2 = torch.zeros_like(smask)
    recs = []
    for n in range(w.shape[0]):
        h1, h2 = rho_habc_step(c, rho, h1, h2, dt, h, b, multiple=multiple)
        y = h1.clone()
        y += smask * w[n]
        h1 = y
        recs.append(h1[bidx, rz, rx])
    stacked = torch.stack(recs, 0).unsqueeze(2)
    return list(torch.split(stacked, counts, dim=1))


def rho_habc_gradient_ad(c, rho, b, dt, h, wavelet, src, rec, rec_cotangent,
                         dtype=torch.float32, multiple=False):
    """Autograd through the whole variable-density loop for the records'
    cotangent: ((dJ/dc, dJ/drho), records)."""
    ps = [torch.as_tensor(x).to(dtype).clone().requires_grad_(True) for x in (c, rho)]
    syn = rho_habc_forward(c, rho, b, dt, h, wavelet, src, rec, dtype, multiple, params=ps)
    torch.cat(syn, 1)[..., 0].backward(torch.as_tensor(rec_cotangent).to(dtype))
    return tuple(p.grad.detach() for p in ps), [s.detach() for s in syn]


def habc_forward(c, b, dt, h, wavelet, src, rec, dtype=torch.float32, c_param=None,
                 multiple=False):
    """Records of the HABC loop (rnn.py:178-212), list of (nt, nrec_b, 1)."""
    c0, b, dt_t, h_t, w, smask, (bidx, rz, rx), counts = _setup(c, b, dt, h, wavelet, src, rec,
                                                                dtype)
    c = c0 if c_param is None else c_param
    h1 = torch.zeros_like(smask)
    h2 = torch.zeros_like(smask)
    recs = []
    for n in range(w.shape[0]):
        h1, h2 = habc_step(c, h1, h2, dt_t, h_t, b, multiple=multiple)
        y = h1.clone()
        y += smask * w[n]
        h1 = y
        recs.append(h1[bidx, rz, rx])
    stacked = torch.stack(recs, 0).unsqueeze(2)
    return list(torch.split(stacked, counts, dim=1))


def habc_gradient_ad(c, b, dt, h, wavelet, src, rec, rec_cotangent, dtype=torch.float32,
                     multiple=False):
    cp = torch.as_tensor(c).to(dtype).clone().requires_grad_(True)
    syn = habc_forward(c, b, dt, h, wavelet, src, rec, dtype, c_param=cp, multiple=multiple)
    torch.cat(syn, 1)[..., 0].backward(torch.as_tensor(rec_cotangent).to(dtype))
    return cp.grad.detach(), [s.detach() for s in syn]
