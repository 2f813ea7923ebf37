"""Pinning the CPU oracle (no GPU).

The reference ships no golden vectors for this path, so each oracle kernel
is held to what the mathematics fixes: closed forms (linear fields, planes,
circles), invariants (telescoping sums, symmetry, discrete
divergence-freeness after a converged projection), brute force on small
grids, and an independent vectorised NumPy restatement.  Each check is
chosen so that a plausible slip (a flipped sign, a half-cell offset, a
transposed operand, a wrong neighbour) makes it fail.
"""
import math

import numpy as np
import pytest

from bevy_eulerian_fluid_amd import scenes
from oracle.oracle import OracleInstance

DT = 1.0 / 60.0


def fresh(W=32, H=32, proj=None, rein=None, length_unit=10.0, rho=99.7, gravity=(0.0, 9.8)):
    s = scenes.settings(W, H, rho=rho, gravity=gravity, length_unit=length_unit,
                        proj=proj or scenes.projection("gauss_seidel", gs_iterations=20),
                        rein=rein or scenes.reinit("fast_iterative", 10))
    return OracleInstance(s)


def face_open_exact(phi_a, phi_b):
    """Open (phi >= 0) fraction of a segment whose end values are phi_a,
    phi_b with phi linear along it."""
    if phi_a >= 0 and phi_b >= 0:
        return 1.0
    if phi_a < 0 and phi_b < 0:
        return 0.0
    return max(phi_a, phi_b) / abs(phi_a - phi_b)


# ---------------------------------------------------------------------------
# area fractions (utils/area_fraction.wgsl) from a planar solid level set
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("normal", [(1.0, 0.3), (-0.4, 1.0), (0.7, -0.7)])
def test_area_fraction_planar(normal):
    W = H = 24
    o = fresh(W, H)
    nx, ny = normal
    c = 0.37
    phi = lambda x, y: nx * (x - 11.3) + ny * (y - 12.6) + c  # noqa: E731
    ys, xs = np.mgrid[0:H, 0:W]
    o.set("LS_SOLID", phi(xs, ys).astype(np.float32))
    o.run_stage("AREA_FRACTION", DT)
    af = o.get("AREA_FRACTION")
    for y in range(2, H - 2):
        for x in range(2, W - 2):
            want = [
                face_open_exact(phi(x - 0.5, y - 0.5), phi(x - 0.5, y + 0.5)),  # i - 1/2
                face_open_exact(phi(x + 0.5, y - 0.5), phi(x + 0.5, y + 0.5)),  # i + 1/2
                face_open_exact(phi(x - 0.5, y - 0.5), phi(x + 0.5, y - 0.5)),  # j - 1/2
                face_open_exact(phi(x - 0.5, y + 0.5), phi(x + 0.5, y + 0.5)),  # j + 1/2
            ]
            np.testing.assert_allclose(af[y, x], want, atol=2e-5, err_msg=f"cell {x},{y}")


def test_area_fraction_shared_faces_agree():
    """A face seen from its two cells must get the same fraction (the
    pressure operator is then symmetric)."""
    o = fresh(40, 40)
    rng = np.random.default_rng(3)
    o.set("LS_SOLID", rng.standard_normal((40, 40)).astype(np.float32) * 3)
    o.run_stage("AREA_FRACTION", DT)
    af = o.get("AREA_FRACTION")
    assert np.array_equal(af[:, :-1, 1], af[:, 1:, 0])
    assert np.array_equal(af[:-1, :, 3], af[1:, :, 2])


# ---------------------------------------------------------------------------
# update_solid (update_solid.wgsl): circle SDF, ids, rigid velocities
# ---------------------------------------------------------------------------
def test_update_solid_circle_closed_form():
    W = H = 64
    o = fresh(W, H)
    cx, cy, r = 5.0, -3.0, 9.0
    omega, vlin = 0.8, (2.0, -1.5)
    ob = scenes.obstacle(7, 0, [r], 0.0, (cx, cy), lin_vel=vlin, ang_vel=omega)
    o.set_obstacles([ob])
    o.run_stage("UPDATE_SOLID", DT)
    ls = o.get("LS_SOLID")
    sid = o.get("SOLID_ID")
    us, vs = o.get("U_SOLID"), o.get("V_SOLID")
    for y in range(1, H - 1):
        for x in range(1, W - 1):
            wx, wy = x - W / 2, H / 2 - y  # world: origin at the centre, y up
            d = math.hypot(wx - cx, wy - cy) - r
            wall = min(x, W - 1 - x, y, H - 1 - y)
            assert ls[y, x] == pytest.approx(min(d, wall), abs=1e-4)
            assert (sid[y, x] == 7) == (d < 0.5)
            # x-face at (x - 1/2, y): u = vx - omega * ry
            fx, fy = x - 0.5 - W / 2, H / 2 - y
            if math.hypot(fx - cx, fy - cy) - r < 0.5:
                assert us[y, x] == pytest.approx(vlin[0] - omega * (fy - cy), abs=1e-4)
            else:
                assert us[y, x] == 0.0
            # y-face at (x, y - 1/2): v = -(vy + omega * rx)  (grid y points down)
            gx, gy = x - W / 2, H / 2 - (y - 0.5)
            if math.hypot(gx - cx, gy - cy) - r < 0.5:
                assert vs[y, x] == pytest.approx(-(vlin[1] + omega * (gx - cx)), abs=1e-4)
    assert (ls[0, :] == 0).all() and (ls[:, -1] == 0).all() and (sid[0, :] == -1).all()


@pytest.mark.parametrize("shape", ["rectangle", "capsule", "triangle"])
def test_update_solid_shapes_brute_force(shape):
    """Rectangle / capsule / triangle SDFs against a dense point-sampled
    distance to the rotated shape's boundary."""
    W = H = 48
    o = fresh(W, H)
    ang = 0.6
    if shape == "rectangle":
        ob = scenes.obstacle(0, 1, [7.0, 4.0], ang, (2.0, 1.0))
        t = np.linspace(-1, 1, 801)
        local = np.concatenate([np.stack([7 * t, 4 * np.ones_like(t)], 1), np.stack([7 * t, -4 * np.ones_like(t)], 1),
                                np.stack([7 * np.ones_like(t), 4 * t], 1), np.stack([-7 * np.ones_like(t), 4 * t], 1)])
    elif shape == "capsule":
        ob = scenes.obstacle(0, 2, [0.0, -6.0, 0.0, 6.0, 3.0], ang, (2.0, 1.0))
        t = np.linspace(0, np.pi, 1000)
        seg = np.linspace(-6, 6, 601)
        caps = [np.stack([3 * np.cos(t), 6 + 3 * np.sin(t)], 1), np.stack([3 * np.cos(t), -6 - 3 * np.sin(t)], 1)]
        local = np.concatenate(caps + [np.stack([3 * np.ones_like(seg), seg], 1),
                                       np.stack([-3 * np.ones_like(seg), seg], 1)])
    else:
        pts = np.array([[-6.0, -4.0], [7.0, -3.0], [0.0, 8.0]])
        ob = scenes.obstacle(0, 4, pts.reshape(-1).tolist(), ang, (2.0, 1.0))
        t = np.linspace(0, 1, 800)[:, None]
        local = np.concatenate([pts[i] + t * (pts[(i + 1) % 3] - pts[i]) for i in range(3)])
    c, s = math.cos(ang), math.sin(ang)
    world = local @ np.array([[c, s], [-s, c]]) + [2.0, 1.0]
    o.set_obstacles([ob])
    o.run_stage("UPDATE_SOLID", DT)
    ls = o.get("LS_SOLID")
    for y in range(8, H - 8, 3):
        for x in range(8, W - 8, 3):
            p = np.array([x - W / 2, H / 2 - y])
            d = np.min(np.hypot(*(world - p).T))
            wall = min(x, W - 1 - x, y, H - 1 - y)
            got = ls[y, x]
            if got < wall - 1e-3:
                assert abs(abs(got) - d) < 0.05, (x, y, got, d)


# ---------------------------------------------------------------------------
# semi-Lagrangian advection (advect_velocity.wgsl, utils/coordinate.wgsl)
# ---------------------------------------------------------------------------
def test_advect_u_linear_field_closed_form():
    """u = a + b * x on x-faces (x = i - 1/2), v = 0: RK2 departure point
    and linear interpolation are exact, so the new u is a + b * x_d."""
    W, H = 40, 64  # x-major dispatch covers 64 * floor(H / 64) rows
    o = fresh(W, H)
    a, b, dt = 3.0, 0.25, 0.05
    xs = np.arange(W + 1) - 0.5
    u = np.tile((a + b * xs).astype(np.float32), (H, 1))
    o.set("U0", u)
    o.set("V0", np.zeros((H + 1, W), np.float32))
    o.run_stage("ADVECT_VELOCITY", dt)
    got = o.get("U1")
    for i in range(3, W - 3):
        x = i - 0.5
        xm = x - 0.5 * dt * (a + b * x)
        xd = x - dt * (a + b * xm)
        np.testing.assert_allclose(got[5:H - 5, i], a + b * xd, rtol=2e-6)


def test_advect_v_linear_field_closed_form():
    W, H = 64, 40  # y-major dispatch covers 64 * floor(W / 64) columns
    o = fresh(W, H)
    a, b, dt = -2.0, 0.3, 0.05
    ys = np.arange(H + 1) - 0.5
    v = np.tile((a + b * ys).astype(np.float32)[:, None], (1, W))
    o.set("V0", v)
    o.set("U0", np.zeros((H, W + 1), np.float32))
    o.run_stage("ADVECT_VELOCITY", dt)
    got = o.get("V1")
    for j in range(3, H - 3):
        y = j - 0.5
        ym = y - 0.5 * dt * (a + b * y)
        yd = y - dt * (a + b * ym)
        np.testing.assert_allclose(got[j, 5:W - 5], a + b * yd, rtol=2e-6)


def test_advect_levelset_uniform_flow_exact_on_quadratics():
    """Catmull-Rom reproduces quadratics; with a uniform velocity (c, d) the
    advected level set is phi(x - c dt, y - d dt)."""
    W = H = 32
    o = fresh(W, H)
    c, d, dt = 7.0, -4.0, 0.1
    o.set("U0", np.full((H, W + 1), c, np.float32))
    o.set("V0", np.full((H + 1, W), d, np.float32))
    phi = lambda x, y: 0.02 * (x - 10) ** 2 + 0.5 * y - 3.0 + 0.01 * x * y  # noqa: E731
    ys, xs = np.mgrid[0:H, 0:W]
    o.set("LS_AIR0", phi(xs, ys).astype(np.float32))
    o.set("LS_AIR1", np.zeros((H, W), np.float32))
    o.run_stage("ADVECT_LEVELSET", dt)
    got = o.get("LS_AIR1")
    want = phi(xs - c * dt, ys - d * dt)
    np.testing.assert_allclose(got[4:-4, 4:-4], want[4:-4, 4:-4], atol=2e-5)


# ---------------------------------------------------------------------------
# divergence.wgsl: sign and scale on a linear field, solid terms
# ---------------------------------------------------------------------------
def test_divergence_linear_field():
    W = H = 24
    o = fresh(W, H, length_unit=4.0)  # dx = 0.25
    o.set("LS_SOLID", np.full((H, W), 50.0, np.float32))  # no solid anywhere
    o.run_stage("AREA_FRACTION", DT)
    xs = np.arange(W + 1, dtype=np.float32)
    ys = np.arange(H + 1, dtype=np.float32)
    o.set("U1", np.tile(2.0 * xs, (H, 1)))
    o.set("V1", np.tile((-0.5 * ys)[:, None], (1, W)))
    o.run_stage("DIVERGENCE", DT)
    div = o.get("DIV")
    np.testing.assert_allclose(div[2:-2, 2:-2], -(2.0 - 0.5) / 0.25, rtol=1e-6)


def test_divergence_closed_solid_faces_use_solid_velocity():
    """With every face closed (fraction 0) only the solid velocities count."""
    W = H = 16
    o = fresh(W, H)
    o.set("LS_SOLID", np.full((H, W), -5.0, np.float32))
    o.run_stage("AREA_FRACTION", DT)
    rng = np.random.default_rng(0)
    us = rng.standard_normal((H, W + 1)).astype(np.float32)
    vs = rng.standard_normal((H + 1, W)).astype(np.float32)
    o.set("U_SOLID", us)
    o.set("V_SOLID", vs)
    o.set("U1", rng.standard_normal((H, W + 1)).astype(np.float32) * 100)
    o.run_stage("DIVERGENCE", DT)
    div = o.get("DIV")
    want = -((us[:, 1:] - us[:, :-1]) + (vs[1:, :] - vs[:-1, :])) / 0.1
    np.testing.assert_allclose(div[2:-2, 2:-2], want[2:-2, 2:-2], rtol=1e-5, atol=1e-4)


# ---------------------------------------------------------------------------
# pressure: GS / MG / MGPCG against the FP64 solve of the same system
# ---------------------------------------------------------------------------
def _tank(W, H, proj, bodies=True):
    from tests.parity_util import small_scene
    sc = small_scene(W, H, proj=proj, bodies=bodies, forces=False)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.set_obstacles(sc["obstacles"])
    o.step(DT)
    for _ in range(3):
        o.step(DT)
    for st in ["UPDATE_SOLID", "AREA_FRACTION", "ADVECT_VELOCITY", "APPLY_FORCES", "DIVERGENCE"]:
        o.run_stage(st, DT)
    return o


@pytest.mark.parametrize("method", ["gauss_seidel", "mgpcg"])
def test_projection_matches_fp64_solve(method):
    W = H = 64
    if method == "gauss_seidel":
        proj = scenes.projection("gauss_seidel", gs_iterations=4000)
    else:
        proj = scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-6)
    o = _tank(W, H, proj)
    o.set("P0", np.zeros((H, W), np.float32))
    p64, rel64, it = o.solve_fp64(DT, 1e-12)
    assert rel64 < 1e-11
    o.run_stage("PROJECTION", DT)
    p = o.get("P0").astype(np.float64)
    ls = o.get("LS_AIR0")
    af = o.get("AREA_FRACTION")
    act = (ls < 0) & ~(af == 0).all(-1)
    scale = np.abs(p64[act]).max()
    assert np.abs(p[act] - p64[act]).max() < 2e-4 * scale


def test_residual_operator_is_linear_and_symmetric():
    """r(x) = b - A x with A symmetric on active cells: <A x, y> = <x, A y>."""
    W = H = 48
    o = _tank(W, H, scenes.projection("gauss_seidel"))
    rng = np.random.default_rng(5)
    ls, af = o.get("LS_AIR0"), o.get("AREA_FRACTION")
    act = (ls < 0) & ~(af == 0).all(-1)
    x = np.where(act, rng.standard_normal((H, W)), 0).astype(np.float32)
    y = np.where(act, rng.standard_normal((H, W)), 0).astype(np.float32)
    b = o.get("DIV")
    Ax = (b - o.residual_of(DT, x)).astype(np.float64)
    Ay = (b - o.residual_of(DT, y)).astype(np.float64)
    lhs, rhs = np.sum(Ax[act] * y[act]), np.sum(x[act] * Ay[act])
    assert abs(lhs - rhs) < 1e-4 * (abs(lhs) + abs(rhs))
    # positive definite on random vectors
    assert np.sum(Ax[act] * x[act]) > 0


def test_hydrostatic_pressure_linear_in_depth():
    """Resting water in a tank: the converged pressure grows by rho * g per
    cell of depth (velocities are in cells/s, so p carries a 1/dx)."""
    W = H = 64
    proj = scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-7)
    s = scenes.settings(W, H, proj=proj, rein=scenes.reinit("fast_iterative", 10))
    o = OracleInstance(s)
    # water from grid row 30 down through the bottom wall, flat surface
    o.set_sources([scenes.source((0.0, -16.0), "aabb", half_size=(40.0, 18.0))], True)
    o.step(DT)
    o.step(DT)
    p = o.get("P0")
    ls = o.get("LS_AIR0")
    col = W // 2
    rows = [y for y in range(36, 56) if ls[y, col] < -2]
    dp = np.diff(p[rows, col].astype(np.float64))
    # one step of gravity: v += g dt/dx (cells/s); solve_velocity removes it
    # with dv = dt/(dx rho) dp, so dp = rho * g per cell of depth
    np.testing.assert_allclose(dp, 99.7 * 9.8, rtol=2e-5)


def test_projection_makes_velocity_divergence_free():
    """D_after = -dx * r: after a tight MGPCG solve the normalised divergence
    of u0/v0 is below 1e-5."""
    W = H = 64
    proj = scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-6)
    from tests.parity_util import small_scene
    sc = small_scene(W, H, proj=proj, forces=False)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.set_obstacles(sc["obstacles"])
    o.set_diagnostics(True)
    for _ in range(4):
        o.step(DT)
    st = o.stats()
    assert st.converged == 1
    assert st.max_norm_divergence < 1e-4, st.max_norm_divergence


# ---------------------------------------------------------------------------
# redistancing
# ---------------------------------------------------------------------------
def test_jfa_matches_brute_force_nearest_seed():
    W, H = 48, 40
    o = fresh(W, H, rein=scenes.reinit("jump_flooding"))
    ys, xs = np.mgrid[0:H, 0:W]
    phi = np.hypot(xs - 20.3, ys - 17.8) - 9.1
    phi = np.minimum(phi, np.hypot(xs - 37.0, ys - 30.0) - 5.2)
    o.set("LS_AIR1", phi.astype(np.float32))
    o.run_stage("REINITIALIZE", DT)
    seeds0 = o.get("JFA_SEEDS0")
    out = o.get("LS_AIR0")
    # the init seeds: brute-force nearest valid seed for every cell
    # (re-derive the seeds independently from the sub-cell crossings)
    pts = []
    for y in range(H):
        for x in range(W):
            if phi[y, x] >= 0:
                continue
            best, bs = 10.0, None
            for i in (-1, 0, 1):
                for j in (-1, 0, 1):
                    if (i, j) == (0, 0) or not (0 <= x + i < W and 0 <= y + j < H):
                        continue
                    n = phi[y + j, x + i]
                    if n >= 0:
                        dd = phi[y, x] / (phi[y, x] - n)
                        if abs(dd) < best:
                            best, bs = abs(dd), (x + dd * i, y + dd * j)
            if bs is not None:
                pts.append(bs)
    pts = np.array(pts)
    d = np.min(np.hypot(xs[..., None] - pts[:, 0], ys[..., None] - pts[:, 1]), axis=-1)
    want = np.where(phi < 0, -d, d)
    err = np.abs(out - want)
    # jump flooding is approximate: a small share of cells keeps a seed that
    # is not the nearest one, by well under a cell
    assert np.mean(err < 1e-4) > 0.97
    assert err.max() < 1.5
    # signed distance property: |grad phi| ~ 1 away from the seeds
    gx = 0.5 * (out[2:-2, 3:-1] - out[2:-2, 1:-3])
    assert np.median(np.abs(gx)) <= 1.0 + 1e-3


def test_fim_planar_interface_exact():
    """A planar interface along x: the eikonal update is exact, so after k
    iterations the band holds |y - y0| exactly for cells within reach."""
    W, H = 32, 40
    o = fresh(W, H, rein=scenes.reinit("fast_iterative", 10))
    y0 = 19.4
    ys, xs = np.mgrid[0:H, 0:W]
    phi = (ys - y0) * 3.0  # not a distance yet (slope 3)
    o.set("LS_AIR1", phi.astype(np.float32))
    o.run_stage("REINITIALIZE", DT)
    out = o.get("LS_AIR0")
    # sources keep their value (rows 19, 20); the band rows 21..27 and 12..18
    np.testing.assert_array_equal(out[19], phi[19].astype(np.float32))
    for y in range(21, 28):
        np.testing.assert_allclose(out[y], out[20, 0] + (y - 20), rtol=1e-6)
    for y in range(12, 19):
        np.testing.assert_allclose(out[y], out[19, 0] - (19 - y), rtol=1e-6)
    # far field keeps the sign only
    assert (out[35] == 1e30).all() and (out[2] == -1e30).all()


def test_fim_circle_distance_error_small():
    W = H = 48
    o = fresh(W, H, rein=scenes.reinit("fast_iterative", 10))
    ys, xs = np.mgrid[0:H, 0:W]
    d = np.hypot(xs - 23.6, ys - 24.2) - 11.3
    o.set("LS_AIR1", (d * 0.5).astype(np.float32))  # wrong slope on purpose
    o.run_stage("REINITIALIZE", DT)
    out = o.get("LS_AIR0")
    band = (np.abs(d) > 1.0) & (np.abs(d) < 6.0)
    assert np.max(np.abs(out[band] - d[band])) < 0.6
    assert np.all(np.sign(out[band]) == np.sign(d[band]))


# ---------------------------------------------------------------------------
# extrapolation, solve_velocity, forces, sources
# ---------------------------------------------------------------------------
def test_extrapolate_constant_fills_air_faces():
    W, H = 64, 64
    o = fresh(W, H)
    ys, xs = np.mgrid[0:H, 0:W]
    o.set("LS_AIR0", (30.5 - ys).astype(np.float32))  # fluid below row 30 ... rows > 30.5
    u = np.where(np.arange(H)[:, None] >= 31, 5.0, 123.0) * np.ones((H, W + 1))
    o.set("U0", u.astype(np.float32))
    o.set("V0", np.zeros((H + 1, W), np.float32))
    o.run_stage("EXTRAPOLATE_VELOCITY", DT)
    got = o.get("U0")
    # faces within 10 layers of the valid region become 5 (the valid value)
    assert np.all(got[22:, 2:-2] == 5.0)
    val = o.get("U_VALID0")
    assert np.all(val[22:, 2:-2] == 1)


def test_solve_velocity_hydrostatic_cancels_gravity():
    """A pressure rising by rho g dx per row cancels one step of gravity:
    v0 = v1 - dt/(dx rho) (p_j - p_{j-1}) = 0 inside the fluid."""
    W = H = 32
    o = fresh(W, H)
    o.set("LS_SOLID", np.full((H, W), 50.0, np.float32))
    o.run_stage("AREA_FRACTION", DT)
    o.set("LS_AIR0", np.full((H, W), -5.0, np.float32))
    g = 9.8
    v1 = np.full((H + 1, W), g * DT / 0.1, np.float32)
    o.set("V1", v1)
    p = (np.arange(H)[:, None] * 99.7 * g * 0.1 * np.ones((1, W))).astype(np.float32)
    o.set("P0", p)
    o.run_stage("SOLVE_VELOCITY", DT)
    v0 = o.get("V0")
    np.testing.assert_allclose(v0[2:H - 2, :], 0.0, atol=1e-3)


def test_apply_forces_gravity_and_gaussian():
    W = H = 64
    o = fresh(W, H, gravity=(1.0, 9.8))
    o.set("LS_SOLID", np.full((H, W), 50.0, np.float32))
    o.run_stage("AREA_FRACTION", DT)
    o.set("LS_AIR0", np.full((H, W), -5.0, np.float32))
    o.set("U1", np.zeros((H, W + 1), np.float32))
    o.set("V1", np.zeros((H + 1, W), np.float32))
    o.set_forces([scenes.force((30.0, -20.0), (20.0, 40.0))])
    o.run_stage("APPLY_FORCES", DT)
    u1, v1 = o.get("U1"), o.get("V1")
    for (y, x) in [(40, 20), (45, 27), (10, 50)]:
        gau = math.exp(-((x - 20.0) ** 2 + (y - 40.0) ** 2) / 200.0)
        assert u1[y, x] == pytest.approx((1.0 + 30.0 * gau) * DT / 0.1, rel=1e-5)
        assert v1[y, x] == pytest.approx((9.8 - 20.0 * gau) * DT / 0.1, rel=1e-5)


def test_dam_break_initial_volume():
    sc = scenes.dam_break(128)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.step(DT)  # initialisation
    ls = o.get("LS_AIR0")
    # AABB centre (0, -25.6), half size (64, 38.4): rows with world y in
    # [-64, 12.8] -> grid rows y >= 64 - 12.8 = 51.2
    want = (np.arange(128)[:, None] >= 51.2) * np.ones((1, 128))
    assert np.mean((ls < 0) == want.astype(bool)) > 0.995


def test_fluid_to_solid_uniform_pressure_zero_force():
    """A closed body in uniform pressure: the face-fraction differences
    telescope to zero net force (up to the fixed-point truncation)."""
    W = H = 64
    o = fresh(W, H)
    ob = scenes.obstacle(0, 0, [10.0], 0.0, (0.0, 0.0))
    o.set_obstacles([ob])
    o.run_stage("UPDATE_SOLID", DT)
    o.set("P0", np.full((H, W), 37.0, np.float32))
    o.run_stage("FLUID_TO_SOLID", DT)
    f = o.get("SOLID_FORCES")
    bins = o.get("SOLID_BINS")
    n_cells = int((o.get("SOLID_ID") == 0).sum())
    assert abs(f[0, 0, 0]) <= 0.1 * n_cells * 0.5 and abs(f[0, 0, 1]) <= 0.1 * n_cells * 0.5
    assert np.array_equal(bins[:, 1:], np.zeros_like(bins[:, 1:]))


def test_step_volume_drift_small():
    """Free-surface runs keep the fluid cell count within a few percent."""
    sc = scenes.dam_break(96)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.step(DT)
    v0 = (o.get("LS_AIR0") < 0).sum()
    for _ in range(40):
        o.step(DT)
    v1 = (o.get("LS_AIR0") < 0).sum()
    assert abs(v1 - v0) / v0 < 0.05
    for f in ["U0", "V0", "P0", "LS_AIR0"]:
        assert np.isfinite(o.get(f)).all()
