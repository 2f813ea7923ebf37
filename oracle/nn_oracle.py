"""CPU oracle of the learned preconditioner (test infrastructure only).

A NumPy restatement of the network libefl runs on MFMA (csrc/efl_nn.hip):
a 4-level U-Net of 3x3 convolutions (zero padding), 2x2 average pooling,
nearest upsampling and channel concatenation, with the same layer table,
parameter layout (per layer W[cout][cin][3][3] then bias[cout]) and rounding
points: weights, the input tensor and every layer output are rounded to BF16
(round to nearest even); sums are exact here (float64) where the GPU
accumulates in FP32.  The backward pass is float64 through the rounded
forward values; with round=False the whole network is float64 and the
gradients can be checked by finite differences.

No counterpart in the reference (which has no learned component).  Used only
by tests/ to pin the GPU kernels: forward outputs and parameter gradients.
"""
import numpy as np

# ca, cb, cout, level, upa, relu, in_a, in_b (tensor ids: 0..14 layer
# outputs, 15 the input, 16..18 the pooled skips of layers 1, 3, 5)
NET = [
    (16, 0, 16, 0, 0, 1, 15, -1),
    (16, 0, 16, 0, 0, 1, 0, -1),
    (16, 0, 32, 1, 0, 1, 16, -1),
    (32, 0, 32, 1, 0, 1, 2, -1),
    (32, 0, 64, 2, 0, 1, 17, -1),
    (64, 0, 64, 2, 0, 1, 4, -1),
    (64, 0, 64, 3, 0, 1, 18, -1),
    (64, 0, 64, 3, 0, 1, 6, -1),
    (64, 64, 64, 2, 1, 1, 7, 5),
    (64, 0, 32, 2, 0, 1, 8, -1),
    (32, 32, 32, 1, 1, 1, 9, 3),
    (32, 0, 16, 1, 0, 1, 10, -1),
    (16, 16, 16, 0, 1, 1, 11, 1),
    (16, 0, 16, 0, 0, 1, 12, -1),
    (16, 0, 16, 0, 0, 0, 13, -1),
]
POOL_SRC = {16: 1, 17: 3, 18: 5}
IN_CH = 16
LS_SCALE = 0.125


def param_count():
    return sum(co * (ca + cb) * 9 + co for ca, cb, co, *_ in NET)


def split_params(p):
    """[(W (cout, cin, 3, 3), b (cout,))] per layer."""
    out, off = [], 0
    for ca, cb, co, *_ in NET:
        n = co * (ca + cb) * 9
        out.append((p[off:off + n].reshape(co, ca + cb, 3, 3), p[off + n:off + n + co]))
        off += n + co
    return out


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3x3(x, w, b):
    """x (N, H, W, Cin), w (Cout, Cin, 3, 3), zero padding; exact sums."""
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((n, h, wd, w.shape[0]))
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("nhwc,oc->nhwo", xp[:, ky:ky + h, kx:kx + wd, :], w[:, :, ky, kx], optimize=True)
    return out + b


def conv3x3_bwd(x, w, dz):
    """Gradients of conv3x3 w.r.t. x, w, b given dz (N, H, W, Cout)."""
    n, h, wd, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for ky in range(3):
        for kx in range(3):
            dw[:, :, ky, kx] = np.einsum("nhwo,nhwc->oc", dz, xp[:, ky:ky + h, kx:kx + wd, :], optimize=True)
            dxp[:, ky:ky + h, kx:kx + wd, :] += np.einsum("nhwo,oc->nhwc", dz, w[:, :, ky, kx], optimize=True)
    return dxp[:, 1:-1, 1:-1, :], dw, dz.sum(axis=(0, 1, 2))


def pool2(x):
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def input_tensor(planes):
    """planes (N, 5, H, W) dataset samples -> (N, H, W, 16) input."""
    n, _, h, w = planes.shape
    x = np.zeros((n, h, w, IN_CH))
    for c in range(3):
        x[..., c] = planes[:, c]
    return x


def forward(params, x, round_=True):
    """All tensors of a forward pass (dict id -> array)."""
    R = bf16 if round_ else (lambda a: np.asarray(a, dtype=np.float64))
    layers = split_params(np.asarray(params, dtype=np.float32 if round_ else np.float64))
    t = {15: R(x)}
    cache = []
    for k, (ca, cb, co, lvl, upa, relu, ia, ib) in enumerate(NET):
        if k in (2, 4, 6):
            pid = 16 + (k - 2) // 2
            t[pid] = R(pool2(t[POOL_SRC[pid]]))
        a = t[ia]
        if upa:
            a = up2(a)
        inp = np.concatenate([a, t[ib]], axis=-1) if ib >= 0 else a
        w, b = layers[k]
        z = conv3x3(inp, R(w), b.astype(np.float64))
        t[k] = R(np.maximum(z, 0.0) if relu else z)
        cache.append(inp)
    return t, cache


def _backward(params, t, cache, g14, round_):
    """Parameter gradients (flat, float64) of a forward pass given the
    gradient of the output tensor (N, H, W, 16)."""
    R = bf16 if round_ else (lambda a: np.asarray(a, dtype=np.float64))
    layers = split_params(np.asarray(params, dtype=np.float32 if round_ else np.float64))
    g = {k: None for k in range(19)}
    g[14] = g14
    grads = []
    for k in range(len(NET) - 1, -1, -1):
        ca, cb, co, lvl, upa, relu, ia, ib = NET[k]
        dz = g[k] * (t[k] > 0) if relu else g[k]
        wk, _ = layers[k]
        dinp, dw, db = conv3x3_bwd(cache[k], R(wk), dz)
        grads.append((k, dw, db))
        if k == 0:
            break
        da = dinp[..., :ca]
        if upa:
            da = da[:, 0::2, 0::2] + da[:, 0::2, 1::2] + da[:, 1::2, 0::2] + da[:, 1::2, 1::2]
        g[ia] = da if g[ia] is None else g[ia] + da
        if ib >= 0:
            g[ib] = dinp[..., ca:] if g[ib] is None else g[ib] + dinp[..., ca:]
        if k in (2, 4, 6):
            pid = 16 + (k - 2) // 2
            src = POOL_SRC[pid]
            dp = 0.25 * up2(g[pid])
            g[src] = dp if g[src] is None else g[src] + dp
    flat = np.zeros(param_count())
    off = 0
    per = {k: (dw, db) for k, dw, db in grads}
    for k, (ca, cb, co, *_r) in enumerate(NET):
        nw = co * (ca + cb) * 9
        dw, db = per[k]
        flat[off:off + nw] = dw.reshape(-1)
        flat[off + nw:off + nw + co] = db
        off += nw + co
    return flat


def loss_and_grads(params, planes, round_=True):
    """Loss = sum mask (o - target)^2 / (N H W) and d loss / d params (float64)."""
    x = input_tensor(planes)
    t, cache = forward(params, x, round_)
    n, h, w = planes.shape[0], planes.shape[2], planes.shape[3]
    count = n * h * w
    target, mask = planes[:, 4].astype(np.float64), planes[:, 3]
    d = np.where(mask != 0, t[14][..., 0] - target, 0.0)
    loss = float((d * d).sum() / count)
    g14 = np.zeros_like(t[14])
    g14[..., 0] = 2.0 * d / count
    return loss, _backward(params, t, cache, g14, round_), t[14][..., 0]


def stencil(ls, afs):
    """The pressure operator of a grid, up to its positive factor dt / (rho
    dx^2), as (diag, off) with off[..., k] the weight of neighbour k (-x, +x,
    -y, +y): (A y)_c = diag_c y_c - sum_k off_k y_k.  residual.wgsl:12-57's
    terms: a fluid neighbour (phi < 0) couples with the face's area fraction
    f_k, an air neighbour adds f_k (1 - phi_k / phi_c) to the diagonal, a
    neighbour outside the grid adds nothing; cells that are not unknowns
    (phi >= 0, or all four fractions zero) have zero rows.  float64."""
    ls = np.asarray(ls, dtype=np.float64)
    afs = np.asarray(afs, dtype=np.float64)
    h, w = ls.shape
    act = (ls < 0) & ~(afs == 0).all(axis=-1)
    diag = np.zeros((h, w))
    off = np.zeros((h, w, 4))
    pad = np.pad(ls, 1, constant_values=np.nan)
    nb = [pad[1:-1, :-2], pad[1:-1, 2:], pad[:-2, 1:-1], pad[2:, 1:-1]]
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(4):
            inside = ~np.isnan(nb[k])
            fluid = inside & (nb[k] < 0)
            air = inside & ~(nb[k] < 0)
            off[..., k] = np.where(act & fluid, afs[..., k], 0.0)
            diag += np.where(act & fluid, afs[..., k], 0.0)
            diag += np.where(act & air, afs[..., k] * (1.0 - nb[k] / ls), 0.0)
    return diag, off


def apply_stencil(diag, off, y):
    """A y for (diag, off) of stencil(); y (..., H, W)."""
    out = diag * y
    out[..., :, 1:] -= off[..., :, 1:, 0] * y[..., :, :-1]
    out[..., :, :-1] -= off[..., :, :-1, 1] * y[..., :, 1:]
    out[..., 1:, :] -= off[..., 1:, :, 2] * y[..., :-1, :]
    out[..., :-1, :] -= off[..., :-1, :, 3] * y[..., 1:, :]
    return out


def residual_loss(y, r, diag, off, mask):
    """Per sample 1 - (r . A y)^2 / (|r|^2 |A y|^2) (r on the unknowns) and
    d loss / d y, summed over the samples and divided by their count."""
    n = y.shape[0]
    r = np.where(mask != 0, r, 0.0)
    ay = apply_stencil(diag, off, y)
    loss, gy = 0.0, np.zeros_like(y)
    for i in range(n):
        rr, ra, aa = (r[i] * r[i]).sum(), (r[i] * ay[i]).sum(), (ay[i] * ay[i]).sum()
        if rr > 0 and aa > 0:
            loss += (1.0 - ra * ra / (rr * aa)) / n
            v = np.where(mask[i] != 0, 2.0 * ra / (rr * aa) * (ra / aa * ay[i] - r[i]) / n, 0.0)
            gy[i] = np.where(mask[i] != 0, apply_stencil(diag[i], off[i], v), 0.0)
        else:
            loss += 1.0 / n
    return loss, gy


def residual_loss_and_grads(params, planes, diag, off, round_=True):
    """The residual-norm loss of the network's outputs on samples planes
    (N, 5, H, W) with stencils (diag (N, H, W), off (N, H, W, 4)), and
    d loss / d params (float64)."""
    x = input_tensor(planes)
    t, cache = forward(params, x, round_)
    y = t[14][..., 0]
    loss, gy = residual_loss(y, planes[:, 0].astype(np.float64), diag, off, planes[:, 3])
    g14 = np.zeros_like(t[14])
    g14[..., 0] = gy
    return loss, _backward(params, t, cache, g14, round_), y
