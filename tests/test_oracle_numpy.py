"""An independent, vectorised NumPy restatement of the stencil kernels of
the pressure path (area fractions, divergence, residual, red-black GS,
restriction, prolongation), compared with the C++ oracle bit for bit on
random inputs.  The two are written separately (loops over cells in C++,
whole-array slicing here), so an indexing slip in either shows up."""
import numpy as np
import pytest

from bevy_eulerian_fluid_amd import scenes
from oracle.oracle import OracleInstance

f32 = np.float32
DT = 1.0 / 60.0


def pad(a, k=1):
    """Zero halo: out-of-bounds loads read 0."""
    return np.pad(a, ((k, k), (k, k)) + ((0, 0),) * (a.ndim - 2))


def np_area_fractions(ls):
    P = pad(ls)
    H, W = ls.shape
    c = lambda dx, dy: P[1 + dy:1 + dy + H, 1 + dx:1 + dx + W]  # noqa: E731
    q = f32(0.25)
    mm = q * (((c(-1, -1) + c(0, -1)) + c(-1, 0)) + c(0, 0))
    pm = q * (((c(0, -1) + c(1, -1)) + c(0, 0)) + c(1, 0))
    mp = q * (((c(-1, 0) + c(0, 0)) + c(-1, 1)) + c(0, 1))
    pp = q * (((c(0, 0) + c(1, 0)) + c(0, 1)) + c(1, 1))

    def fr(a, b):
        with np.errstate(divide="ignore", invalid="ignore"):
            v = np.maximum(a, b) / np.abs(a - b)
        v = np.where(np.isnan(v), f32(0), v)
        return np.minimum(np.maximum(v, f32(0)), f32(1)).astype(f32)

    return np.stack([fr(mm, mp), fr(pm, pp), fr(mm, pm), fr(mp, pp)], -1)


def np_divergence(u1, v1, us, vs, lss, dx):
    f = np_area_fractions(lss)
    du_f = f[..., 1] * u1[:, 1:] - f[..., 0] * u1[:, :-1]
    dv_f = f[..., 3] * v1[1:, :] - f[..., 2] * v1[:-1, :]
    one = f32(1)
    du_s = (one - f[..., 1]) * us[:, 1:] - (one - f[..., 0]) * us[:, :-1]
    dv_s = (one - f[..., 3]) * vs[1:, :] - (one - f[..., 2]) * vs[:-1, :]
    return (-(((du_f + dv_f) + du_s) + dv_s) / f32(dx)).astype(f32)


def neighbours(a):
    """(left, right, up, down) with an in-bounds mask, for a 2D array."""
    H, W = a.shape
    P = pad(a)
    nb = [P[1:H + 1, 0:W], P[1:H + 1, 2:W + 2], P[0:H, 1:W + 1], P[2:H + 2, 1:W + 1]]
    ys, xs = np.mgrid[0:H, 0:W]
    inb = [xs > 0, xs < W - 1, ys > 0, ys < H - 1]
    return nb, inb


def np_residual(x, b, ls, af, dx, dt, rho):
    dxx = f32(dx)
    factor = f32(dt) / f32(rho) / dxx / dxx
    xn, inb = neighbours(x)
    ln, _ = neighbours(ls)
    res = b.copy()
    one = f32(1)
    for k in range(4):
        with np.errstate(divide="ignore", invalid="ignore"):
            fluid = res - af[..., k] * (x - xn[k]) * factor
            ghost = res - af[..., k] * (one - ln[k] / ls) * x * factor
        res = np.where(inb[k], np.where(ln[k] < 0, fluid, ghost), res)
    closed = (af == 0).all(-1)
    return np.where((ls >= 0) | closed, f32(0), res).astype(f32)


def np_gs_colour(x, b, ls, af, parity, weight, dx, dt, rho):
    H, W = x.shape
    dxx = f32(dx)
    xn, inb = neighbours(x)
    ln, _ = neighbours(ls)
    denom = np.zeros_like(x)
    nume = dxx * dxx * f32(rho) / f32(dt) * b
    one = f32(1)
    for k in range(4):
        with np.errstate(divide="ignore", invalid="ignore"):
            fl = ln[k] < 0
            d_new = np.where(fl, denom + af[..., k], denom + af[..., k] * (one - ln[k] / ls))
            n_new = np.where(fl, nume + af[..., k] * xn[k], nume)
        denom = np.where(inb[k], d_new, denom)
        nume = np.where(inb[k], n_new, nume)
    with np.errstate(divide="ignore", invalid="ignore"):
        pn = nume / denom
    w = f32(weight)
    val = w * pn + (one - w) * x
    val = np.where(np.abs(denom) < f32(1e-6), f32(0), val)
    val = np.where((ls >= 0) | (af == 0).all(-1), f32(0), val)
    ys, xs = np.mgrid[0:H, 0:W]
    return np.where(((xs + ys) & 1) == parity, val, x).astype(f32)


def scene_state(W, H, seed):
    from tests.parity_util import small_scene
    sc = small_scene(W, H, seed=seed, forces=False)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.set_sources(sc["sources"], False)
    o.set_obstacles(sc["obstacles"])
    for _ in range(4):
        o.step(DT)
    for st in ["UPDATE_SOLID", "AREA_FRACTION", "ADVECT_VELOCITY", "APPLY_FORCES", "DIVERGENCE"]:
        o.run_stage(st, DT)
    return o


@pytest.mark.parametrize("size", [(64, 64), (48, 40)])
def test_area_fractions_and_divergence_numpy(size):
    o = scene_state(*size, seed=11)
    lss = o.get("LS_SOLID")
    np.testing.assert_array_equal(np_area_fractions(lss), o.get("AREA_FRACTION"))
    div = np_divergence(o.get("U1"), o.get("V1"), o.get("U_SOLID"), o.get("V_SOLID"), lss, 0.1)
    np.testing.assert_array_equal(div, o.get("DIV"))


@pytest.mark.parametrize("size", [(64, 64), (48, 40)])
def test_residual_and_gs_numpy(size):
    W, H = size
    o = scene_state(W, H, seed=12)
    rng = np.random.default_rng(2)
    p = (rng.standard_normal((H, W)) * 100).astype(f32)
    o.set("P0", p)
    b, ls, af = o.get("DIV"), o.get("LS_AIR0"), o.get("AREA_FRACTION")
    o.run_stage("MG_RESIDUAL", DT)
    np.testing.assert_array_equal(np_residual(p, b, ls, af, 0.1, DT, 99.7), o.get("MG_R0"))
    o.run_stage("GS_SWEEP", DT)
    x = np_gs_colour(p, b, ls, af, 1, 1.9, 0.1, DT, 99.7)
    x = np_gs_colour(x, b, ls, af, 0, 1.9, 0.1, DT, 99.7)
    np.testing.assert_array_equal(x, o.get("P0"))


def test_restriction_prolongation_numpy():
    """One reference V-cycle on a 2-level hierarchy equals pre-smooth,
    residual, 4-cell average restriction, coarse smooth, injection and
    post-smooth restated in NumPy."""
    W = H = 32  # log2(32) - 2 = 3 levels
    proj = scenes.projection("multigrid", pre=1, post=1, coarsest=2, max_cycles=1)
    from tests.parity_util import small_scene
    sc = small_scene(W, H, seed=13, proj=proj, forces=False)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.set_obstacles(sc["obstacles"])
    for _ in range(3):
        o.step(DT)
    for st in ["UPDATE_SOLID", "AREA_FRACTION", "ADVECT_VELOCITY", "APPLY_FORCES", "DIVERGENCE"]:
        o.run_stage(st, DT)
    x0 = o.get("P0").copy()
    b0, ls0, af0 = o.get("DIV"), o.get("LS_AIR0"), o.get("AREA_FRACTION")
    o.run_stage("MG_VCYCLE", DT)

    def smooth(x, b, ls, af, scale, n):
        for _ in range(n):
            x = np_gs_colour(x, b, ls, af, 1, 1.0, 0.1 * scale, DT, 99.7)
            x = np_gs_colour(x, b, ls, af, 0, 1.0, 0.1 * scale, DT, 99.7)
        return x

    def restrict(r, ls, af):
        q = f32(0.25)
        b = q * (((r[0::2, 0::2] + r[1::2, 0::2]) + r[0::2, 1::2]) + r[1::2, 1::2])
        cands = [ls[0::2, 0::2], ls[1::2, 0::2], ls[0::2, 1::2], ls[1::2, 1::2]]
        lc = cands[0]
        for c in cands[1:]:
            lc = np.where(np.abs(c) < np.abs(lc), c, lc)
        h = f32(0.5)
        ac = np.stack([h * (af[0::2, 0::2, 0] + af[1::2, 0::2, 0]), h * (af[0::2, 1::2, 1] + af[1::2, 1::2, 1]),
                       h * (af[0::2, 0::2, 2] + af[0::2, 1::2, 2]), h * (af[1::2, 0::2, 3] + af[1::2, 1::2, 3])], -1)
        return b.astype(f32), lc, ac.astype(f32)

    def vcycle(x, b, ls, af, lvl, nlev):
        scale = float(1 << lvl)
        if lvl == nlev - 1:
            return smooth(x, b, ls, af, scale, 2)
        x = smooth(x, b, ls, af, scale, 1)
        r = np_residual(x, b, ls, af, 0.1 * scale, DT, 99.7)
        bc, lc, ac = restrict(r, ls, af)
        xc = vcycle(np.zeros_like(bc), bc, lc, ac, lvl + 1, nlev)
        up = np.repeat(np.repeat(xc, 2, 0), 2, 1)
        x = np.where(ls < 0, up + x, x).astype(f32)
        return smooth(x, b, ls, af, scale, 1)

    want = vcycle(x0, b0, ls0, af0, 0, 3)
    np.testing.assert_array_equal(want, o.get("P0"))
