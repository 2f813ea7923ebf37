"""The CPU oracle of the learned preconditioner (oracle/nn_oracle.py), pinned
without the GPU: its convolution, pooling and upsampling against PyTorch's
float32 operators, its backward pass against central finite differences of
the float64 network, its BF16 rounding against torch's bfloat16 cast, and its
parameter layout against the library's count."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from bevy_eulerian_fluid_amd import native
from oracle import nn_oracle as O


def test_param_count_matches_library():
    assert O.param_count() == native.Network.param_count()
    assert 0.2e6 < O.param_count() < 0.35e6


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(10000) * 10.0 ** rng.integers(-8, 8, 10000)).astype(np.float32)
    x[:4] = [1.00390625, 1.01171875, -2.0, 0.0]  # ties to even
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(O.bf16(x), ref)


def test_conv_pool_up_match_torch():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 12, 10, 16))
    w = rng.standard_normal((8, 16, 3, 3))
    b = rng.standard_normal(8)
    ours = O.conv3x3(x, w, b)
    ref = F.conv2d(torch.from_numpy(x).permute(0, 3, 1, 2).float(), torch.from_numpy(w).float(),
                   torch.from_numpy(b).float(), padding=1).permute(0, 2, 3, 1).double().numpy()
    assert np.allclose(ours, ref, rtol=1e-4, atol=1e-4)
    xt = torch.from_numpy(x).permute(0, 3, 1, 2)
    assert np.allclose(O.pool2(x), F.avg_pool2d(xt, 2).permute(0, 2, 3, 1).numpy())
    assert np.allclose(O.up2(x), F.interpolate(xt, scale_factor=2, mode="nearest").permute(0, 2, 3, 1).numpy())


def _planes(rng, n, h, w):
    p = np.zeros((n, 5, h, w))
    p[:, 0] = rng.standard_normal((n, h, w))
    p[:, 1] = rng.uniform(0, 1, (n, h, w))
    p[:, 2] = rng.uniform(-1, 1, (n, h, w))
    p[:, 3] = rng.uniform(0, 1, (n, h, w)) < 0.7
    p[:, 4] = rng.standard_normal((n, h, w))
    return p


@pytest.mark.parametrize("seed", [3, 4])
def test_backward_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    params = rng.standard_normal(O.param_count()) * 0.1
    planes = _planes(rng, 1, 16, 16)
    loss, grads, _ = O.loss_and_grads(params, planes, round_=False)
    # one weight and one bias of every layer, plus a few random entries
    idx = []
    off = 0
    for ca, cb, co, *_ in O.NET:
        nw = co * (ca + cb) * 9
        idx += [off + int(rng.integers(0, min(nw, 9 * co * 3))), off + nw + int(rng.integers(0, co))]
        off += nw + co
    idx += list(rng.integers(0, O.param_count(), 10))
    eps = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (O.loss_and_grads(pp, planes, False)[0] - O.loss_and_grads(pm, planes, False)[0]) / (2 * eps)
        assert abs(fd - grads[i]) <= 2e-4 * max(1.0, abs(fd)) + 1e-7, (i, fd, grads[i])


def _oracle_state(W=48, H=40, steps=3):
    """An oracle instance after a few steps of the compact scene (fluid,
    free surface, four bodies), its level set and area fractions."""
    from oracle.oracle import OracleInstance
    from tests.parity_util import small_scene
    sc = small_scene(W, H)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.set_sources(sc["sources"], False)
    o.set_obstacles(sc["obstacles"])
    for _ in range(steps):
        o.step(sc["dt"])
    return o, sc


def test_stencil_is_the_oracle_pressure_operator():
    """stencil() applied to a random x equals the oracle's residual operator
    (residual.wgsl:12-57, pinned by the FP64 solve and closed forms): A x =
    (r(0) - r(x)) / (dt / (rho dx^2)), for every cell."""
    o, sc = _oracle_state()
    try:
        ls = o.get("LS_AIR0")
        afs = o.get("AREA_FRACTION")
        diag, off = O.stencil(ls, afs)
        s = sc["settings"]
        dx = 1.0 / s.length_unit  # the grid spacing in metres
        factor = sc["dt"] / s.rho / (dx * dx)
        rng = np.random.default_rng(8)
        act = (ls < 0) & ~(afs == 0).all(axis=-1)
        assert act.sum() > 200 and (~act).sum() > 200
        assert (diag[~act] == 0).all() and (off[~act] == 0).all()
        for _ in range(3):
            x = np.where(act, rng.standard_normal(ls.shape), 0.0).astype(np.float32)
            ax = (o.residual_of(sc["dt"], np.zeros_like(x)).astype(np.float64)
                  - o.residual_of(sc["dt"], x).astype(np.float64)) / factor
            ours = O.apply_stencil(diag, off, x.astype(np.float64))
            assert np.allclose(ours, ax, rtol=1e-4, atol=1e-4 * np.abs(ax).max()), np.abs(ours - ax).max()
        # symmetric on the unknowns: off-diagonal weights agree from both cells
        assert np.allclose(off[:, 1:, 0], off[:, :-1, 1]) and np.allclose(off[1:, :, 2], off[:-1, :, 3])
    finally:
        o.close()


def test_residual_loss_closed_forms():
    """The loss is 0 for any multiple of A^-1 r, 1 for an output whose A y is
    orthogonal to r or zero, and 1 - cos^2 in between."""
    o, sc = _oracle_state()
    try:
        ls, afs = o.get("LS_AIR0"), o.get("AREA_FRACTION")
    finally:
        o.close()
    diag, off = O.stencil(ls, afs)
    act = diag != 0
    ids = np.flatnonzero(act)
    n = ids.size
    A = np.zeros((n, n))
    pos = -np.ones(act.size, int)
    pos[ids] = np.arange(n)
    H, W = ls.shape
    for i, c in enumerate(ids):
        y, x = divmod(c, W)
        A[i, i] = diag[y, x]
        for k, (dx, dy) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
            if off[y, x, k] != 0:
                A[i, pos[(y + dy) * W + x + dx]] -= off[y, x, k]
    rng = np.random.default_rng(9)
    r = np.zeros(act.size)
    r[ids] = rng.standard_normal(n)
    mask = act.astype(float)[None]
    y = np.zeros(act.size)
    y[ids] = np.linalg.solve(A, r[ids])
    for c in (1.0, -3.5, 1e-3):
        loss, _ = O.residual_loss((c * y).reshape(1, H, W), r.reshape(1, H, W), diag[None], off[None], mask)
        assert loss < 1e-10, (c, loss)
    loss, g = O.residual_loss(np.zeros((1, H, W)), r.reshape(1, H, W), diag[None], off[None], mask)
    assert loss == 1.0 and not g.any()
    # A y orthogonal to r: y = A^-1 w with w orthogonal to r
    w = rng.standard_normal(n)
    w -= (w @ r[ids]) / (r[ids] @ r[ids]) * r[ids]
    y2 = np.zeros(act.size)
    y2[ids] = np.linalg.solve(A, w)
    loss, _ = O.residual_loss(y2.reshape(1, H, W), r.reshape(1, H, W), diag[None], off[None], mask)
    assert abs(loss - 1.0) < 1e-10
    y3 = np.zeros(act.size)
    y3[ids] = rng.standard_normal(n)
    ay = A @ y3[ids]
    cos2 = (ay @ r[ids]) ** 2 / ((ay @ ay) * (r[ids] @ r[ids]))
    loss, _ = O.residual_loss(y3.reshape(1, H, W), r.reshape(1, H, W), diag[None], off[None], mask)
    assert abs(loss - (1.0 - cos2)) < 1e-12


@pytest.mark.parametrize("seed", [5])
def test_residual_loss_backward_matches_finite_differences(seed):
    """d loss / d params of the residual-norm loss through the float64
    network against central differences, on a real oracle stencil."""
    o, sc = _oracle_state(16, 16, steps=2)
    try:
        ls, afs = o.get("LS_AIR0"), o.get("AREA_FRACTION")
    finally:
        o.close()
    diag, off = O.stencil(ls, afs)
    rng = np.random.default_rng(seed)
    params = rng.standard_normal(O.param_count()) * 0.1
    planes = _planes(rng, 1, 16, 16)
    planes[:, 3] = diag != 0
    loss, grads, _ = O.residual_loss_and_grads(params, planes, diag[None], off[None], round_=False)
    assert 0.0 < loss < 1.0
    idx = []
    off_p = 0
    for ca, cb, co, *_ in O.NET:
        nw = co * (ca + cb) * 9
        idx += [off_p + int(rng.integers(0, min(nw, 9 * co * 3))), off_p + nw + int(rng.integers(0, co))]
        off_p += nw + co
    eps = 1e-6
    f = lambda p: O.residual_loss_and_grads(p, planes, diag[None], off[None], round_=False)[0]
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (f(pp) - f(pm)) / (2 * eps)
        assert abs(fd - grads[i]) <= 2e-4 * max(1.0, abs(fd)) + 1e-7, (i, fd, grads[i])
