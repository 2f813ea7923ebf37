"""Pinning the CPU oracle, part 2 (no GPU): closed forms for the kernels
that the first set of pins left open.

- levelset_gradient.wgsl:4-20: central differences are exact on quadratics.
- jacobi_iteration.wgsl:18-91: the converged iterate satisfies the
  ghost-fluid equations assembled here in float64, and equals -1/dx times
  the Gauss-Seidel system's solution (the Jacobi kernel scales the
  divergence by dx rho / dt with a minus sign, the Gauss-Seidel kernel by
  +dx^2 rho / dt).
- restriction.wgsl:10-59 / prolongation.wgsl:5-33: a linear residual
  restricts to its value at the coarse cell centre, the level set to the
  smallest |phi| of each 2x2 block (first in the shader's order on ties),
  area fractions to the open share of each coarse face of a planar solid;
  a coarse correction is added to the fluid (phi < 0) fine cells only.
- multigrid repeated to a residual target converges to the FP64 solve.
- update_fluid_source.wgsl:30-69: sinks carve max(phi, -sdf) unless a
  source came first; a source stamps its velocity on both faces of each
  covered cell; the AABB distance in closed form.
- sample_forces.wgsl:20-53 + accumulate_forces.wgsl: for a submerged body
  in hydrostatic pressure the vertical force bin is -rho g A dx (A the
  closed area the area fractions describe, close to the shape's area), the
  torque bin is 0 for a centred circle and -rho g A dx (x_centroid - x_com)
  for a capsule whose centre of mass record is off its centroid.
- the rigid-body integrator (no reference counterpart): first-step velocity
  of a body of density rho_b in resting water is dt (g/dx) (rho - rho_b) /
  (rho_b + 4 rho); a neutrally buoyant body stays at rest.
Each check is chosen so that a flipped sign, a x10 slip, a transposed
operand or a half-cell offset fails it.
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
    """Open (phi >= 0) share of a segment with phi linear from phi_a to phi_b."""
    if phi_a >= 0 and phi_b >= 0:
        return 1.0
    if phi_a < 0 and phi_b < 0:
        return 0.0
    return max(phi_a, phi_b) / abs(phi_a - phi_b)


# ---------------------------------------------------------------------------
# levelset_gradient.wgsl
# ---------------------------------------------------------------------------
def test_levelset_gradient_exact_on_quadratics():
    W, H = 40, 32
    o = fresh(W, H)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    phi = 0.03 * (xs - 13.0) ** 2 - 0.02 * (ys - 7.0) ** 2 + 0.01 * xs * ys + 0.5 * xs - 2.0
    o.set("LS_AIR0", phi.astype(np.float32))
    sentinel = np.full((H, W, 2), 77.0, np.float32)
    o.set("GRAD_LS", sentinel)
    o.run_stage("LEVELSET_GRADIENT", DT)
    g = o.get("GRAD_LS").astype(np.float64)
    dphidx = 0.06 * (xs - 13.0) + 0.01 * ys + 0.5
    dphidy = -0.04 * (ys - 7.0) + 0.01 * xs
    sl = (slice(1, H - 1), slice(1, W - 1))
    np.testing.assert_allclose(g[sl][..., 0], dphidx[sl], atol=2e-5)
    np.testing.assert_allclose(g[sl][..., 1], dphidy[sl], atol=2e-5)
    # row 0 and column 0 are not written (any(idx == 0) returns)
    assert (g[0] == 77.0).all() and (g[:, 0] == 77.0).all()
    # the last column differences against the zero that an out-of-range load returns
    p32 = phi.astype(np.float32)
    np.testing.assert_allclose(g[5, W - 1, 0], 0.5 * (0.0 - p32[5, W - 2]), rtol=1e-6)


# ---------------------------------------------------------------------------
# jacobi_iteration.wgsl
# ---------------------------------------------------------------------------
def _jacobi_scene(W=24, H=24):
    proj = scenes.projection("jacobi")
    o = fresh(W, H, proj=proj)
    o.set_obstacles([scenes.obstacle(0, 0, [3.3], 0.0, (2.4, -4.1))])
    o.run_stage("UPDATE_SOLID", DT)
    o.run_stage("AREA_FRACTION", DT)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    ls = 0.3 * (xs - 11.0) + (8.6 - ys)  # fluid below a tilted free surface
    o.set("LS_AIR0", ls.astype(np.float32))
    rng = np.random.default_rng(11)
    o.set("DIV", rng.uniform(-3.0, 3.0, (H, W)).astype(np.float32))
    o.set("P0", np.zeros((H, W), np.float32))
    o.set("P1", np.zeros((H, W), np.float32))
    return o


def _jacobi_system(o, W, H):
    """The ghost-fluid equations jacobi_iteration.wgsl encodes, assembled in
    float64: coef p_c - sum_k f_k [phi_k/phi_c >= 0] p_k = -(dx rho / dt) div_c
    with coef = sum_k f_k (1 - min(0, phi_k / phi_c)), out-of-range loads 0."""
    ls = np.pad(o.get("LS_AIR0").astype(np.float64), 1)
    af = o.get("AREA_FRACTION").astype(np.float64)
    div = o.get("DIV").astype(np.float64)
    dx, rho = 0.1, 99.7
    factor = dx * rho / DT
    c = ls[1:-1, 1:-1]
    nbr = [ls[1:-1, :-2], ls[1:-1, 2:], ls[:-2, 1:-1], ls[2:, 1:-1]]  # -x, +x, -y, +y
    with np.errstate(divide="ignore", invalid="ignore"):
        ff = [np.minimum(0.0, n / c) for n in nbr]
    coef = sum(af[..., k] * (1.0 - ff[k]) for k in range(4))
    fluid_nbr = [(f >= 0.0).astype(np.float64) for f in ff]
    active = (c < 0) & ~(af == 0).all(-1) & (np.abs(coef) >= 1e-6)
    return coef, fluid_nbr, af, factor, div, active


def test_jacobi_fixed_point_satisfies_ghost_fluid_equations():
    W = H = 24
    o = _jacobi_scene(W, H)
    for _ in range(40):  # 4000 Jacobi sweeps
        o.run_stage("PROJECTION", DT)
    p = o.get("P0").astype(np.float64)
    coef, fl, af, factor, div, act = _jacobi_system(o, W, H)
    pp = np.pad(p, 1)
    nbp = [pp[1:-1, :-2], pp[1:-1, 2:], pp[:-2, 1:-1], pp[2:, 1:-1]]
    res = coef * p - sum(af[..., k] * fl[k] * nbp[k] for k in range(4)) + factor * div
    scale = np.max(np.abs(coef * p)[act])
    assert act.sum() > 150
    assert np.max(np.abs(res[act])) < 2e-5 * scale
    # cells in the air are 0
    assert (p[o.get("LS_AIR0") >= 0] == 0.0).all()


def test_jacobi_solution_is_minus_gs_solution_over_dx():
    """The two kernels share the ghost coefficients f_k (1 - phi_k/phi_c) and
    differ only in the right-hand side: Jacobi's fixed point is -p_GS / dx,
    where p_GS is the FP64 solve of the Gauss-Seidel system."""
    W = H = 24
    o = _jacobi_scene(W, H)
    for _ in range(40):
        o.run_stage("PROJECTION", DT)
    pj = o.get("P0").astype(np.float64)
    p64, rel, _ = o.solve_fp64(DT, 1e-13)
    assert rel < 1e-11
    *_, act = _jacobi_system(o, W, H)
    want = -p64 / 0.1
    scale = np.abs(want[act]).max()
    assert np.abs(pj[act] - want[act]).max() < 1e-4 * scale


# ---------------------------------------------------------------------------
# restriction.wgsl / prolongation.wgsl
# ---------------------------------------------------------------------------
def _mg_instance(W=64, H=48):
    return fresh(W, H, proj=scenes.projection("multigrid"))


def test_restriction_of_linear_residual_is_centre_value():
    o = _mg_instance()
    w, h = o.mg_size(0)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    r = 1.5 + 0.25 * xs - 0.75 * ys
    o.mg_set(0, "r", r)
    o.mg_set(1, "x", np.full(o.mg_size(1)[::-1], 9.0))
    o.mg_restrict(0)
    cw, ch = o.mg_size(1)
    assert (cw, ch) == (w // 2, h // 2)
    cy, cx = np.mgrid[0:ch, 0:cw].astype(np.float64)
    want = 1.5 + 0.25 * (2 * cx + 0.5) - 0.75 * (2 * cy + 0.5)  # value at the 2x2 block's centre
    np.testing.assert_allclose(o.mg_get(1, "b"), want, rtol=1e-6, atol=1e-6)
    assert (o.mg_get(1, "x") == 0.0).all()  # the coarse guess starts at zero


def test_restriction_levelset_takes_smallest_magnitude():
    o = _mg_instance(16, 16)
    ls = np.full((16, 16), 50.0, np.float32)

    def block(bx, by, v00, v01, v10, v11):  # (x, y) offsets: (0,0), (0,1), (1,0), (1,1)
        ls[2 * by, 2 * bx] = v00
        ls[2 * by + 1, 2 * bx] = v01
        ls[2 * by, 2 * bx + 1] = v10
        ls[2 * by + 1, 2 * bx + 1] = v11

    block(0, 0, 3.0, -1.0, 2.0, -0.5)   # -> -0.5
    block(1, 0, -4.0, 6.0, -7.0, 5.0)   # -> -4 (the first cell)
    block(2, 0, 5.0, -1.0, 1.0, 2.0)    # |-1| == |1|: (0,1) comes before (1,0) -> -1
    block(3, 0, 2.0, 2.0, -2.0, -2.0)   # all ties: the first -> 2
    block(0, 1, 0.25, 0.5, -0.125, 9.0)  # -> -0.125
    o.mg_set(0, "ls", ls)
    o.mg_restrict(0)
    c = o.mg_get(1, "ls")
    assert c[0, 0] == np.float32(-0.5)
    assert c[0, 1] == np.float32(-4.0)
    assert c[0, 2] == np.float32(-1.0)
    assert c[0, 3] == np.float32(2.0)
    assert c[1, 0] == np.float32(-0.125)
    assert c[3, 5] == np.float32(50.0)


@pytest.mark.parametrize("normal", [(1.0, 0.35), (-0.3, 1.0), (0.6, -0.8)])
def test_restriction_area_fractions_of_planar_solid(normal):
    """The coarse cell's faces are twice as long; for a planar solid the
    coarse face's open share is the open share of that longer segment."""
    W = H = 32
    o = _mg_instance(W, H)
    nx, ny = normal
    phi = lambda x, y: nx * (x - 15.3) + ny * (y - 16.2) + 0.21  # noqa: E731
    ys, xs = np.mgrid[0:H, 0:W]
    o.set("LS_SOLID", phi(xs, ys).astype(np.float32))
    o.run_stage("AREA_FRACTION", DT)
    o.mg_restrict(0)
    af = o.mg_get(1, "af")
    for Y in range(2, H // 2 - 2):
        for X in range(2, W // 2 - 2):
            x0, x1 = 2 * X - 0.5, 2 * X + 1.5  # coarse cell edges in fine coordinates
            y0, y1 = 2 * Y - 0.5, 2 * Y + 1.5
            want = [face_open_exact(phi(x0, y0), phi(x0, y1)), face_open_exact(phi(x1, y0), phi(x1, y1)),
                    face_open_exact(phi(x0, y0), phi(x1, y0)), face_open_exact(phi(x0, y1), phi(x1, y1))]
            np.testing.assert_allclose(af[Y, X], want, atol=3e-5, err_msg=f"coarse cell {X},{Y}")


def test_prolongation_adds_coarse_value_to_fluid_cells_only():
    o = _mg_instance()
    w, h = o.mg_size(0)
    cw, ch = o.mg_size(1)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((h, w)).astype(np.float32)
    ls = rng.standard_normal((h, w)).astype(np.float32)
    xc = rng.standard_normal((ch, cw)).astype(np.float32)
    o.mg_set(0, "x", x)
    o.mg_set(0, "ls", ls)
    o.mg_set(1, "x", xc)
    o.mg_prolong(0)
    got = o.mg_get(0, "x")
    up = np.repeat(np.repeat(xc, 2, axis=0), 2, axis=1)
    want = np.where(ls < 0, up + x, x)  # one float32 add: exact
    assert np.array_equal(got, want)


def test_multigrid_to_tolerance_matches_fp64_solve():
    from tests.parity_util import small_scene
    W = H = 64
    proj = scenes.projection("multigrid", max_cycles=400, rel_tol=1e-5)  # float32 floor is ~1.2e-6
    sc = small_scene(W, H, proj=proj, forces=False)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.set_obstacles(sc["obstacles"])
    for _ in range(4):
        o.step(DT)
    for st in ["UPDATE_SOLID", "AREA_FRACTION", "ADVECT_VELOCITY", "APPLY_FORCES", "DIVERGENCE"]:
        o.run_stage(st, DT)
    o.set("P0", np.zeros((H, W), np.float32))
    p64, rel64, _ = o.solve_fp64(DT, 1e-12)
    o.run_stage("PROJECTION", DT)
    st = o.stats()
    assert st.converged == 1 and st.rel_residual <= 1e-5 and 1 < st.cycles < 50
    p = o.get("P0").astype(np.float64)
    ls, af = o.get("LS_AIR0"), o.get("AREA_FRACTION")
    act = (ls < 0) & ~(af == 0).all(-1)
    assert np.abs(p[act] - p64[act]).max() < 5e-5 * np.abs(p64[act]).max()


# ---------------------------------------------------------------------------
# update_fluid_source.wgsl
# ---------------------------------------------------------------------------
def _cell_world(W, H):
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    return xs - W / 2, H / 2 - ys  # uv_to_ndc: (uv * (2, -2) + (-1, 1)) * dim / 2


def _source_instance(W, H, sources, ls_value=-5.0):
    o = fresh(W, H)
    o.set("LS_AIR1", np.full((H, W), ls_value, np.float32))
    o.set("U0", np.full((H, W + 1), 123.0, np.float32))
    o.set("V0", np.full((H + 1, W), 123.0, np.float32))
    o.set_sources(sources, False)
    o.run_stage("SOURCE", DT)
    return o


def test_sink_carves_negative_sdf():
    W, H = 48, 40
    c, R = (5.0, -3.0), 7.5
    o = _source_instance(W, H, [scenes.source(c, "circle", radius=R, sink=True)])
    px, py = _cell_world(W, H)
    d = np.hypot(px - c[0], py - c[1]) - R
    want = np.maximum(-5.0, -d)
    np.testing.assert_allclose(o.get("LS_AIR1"), want, atol=2e-5)
    # a sink stamps no velocity
    assert (o.get("U0") == 123.0).all() and (o.get("V0") == 123.0).all()


def test_sink_after_source_is_ignored_sink_before_source_applies():
    W, H = 48, 40
    src = scenes.source((-10.0, 0.0), "circle", radius=4.0)
    snk = scenes.source((8.0, 2.0), "circle", radius=6.0, sink=True)
    px, py = _cell_world(W, H)
    d_src = np.hypot(px + 10.0, py - 0.0) - 4.0
    d_snk = np.hypot(px - 8.0, py - 2.0) - 6.0
    after = _source_instance(W, H, [src, snk]).get("LS_AIR1")  # has_source: the sink is skipped
    np.testing.assert_allclose(after, np.minimum(-5.0, d_src), atol=2e-5)
    before = _source_instance(W, H, [snk, src]).get("LS_AIR1")
    np.testing.assert_allclose(before, np.minimum(np.maximum(-5.0, -d_snk), d_src), atol=2e-5)


def test_source_stamps_velocity_on_both_faces_of_covered_cells():
    W, H = 48, 40
    a = scenes.source((-9.0, 4.0), "circle", radius=5.2, velocity=(3.5, -1.25))
    b = scenes.source((11.0, -6.0), "aabb", half_size=(4.0, 3.0), velocity=(-2.0, 6.0))
    o = _source_instance(W, H, [a, b], ls_value=20.0)
    px, py = _cell_world(W, H)
    in_a = np.hypot(px + 9.0, py - 4.0) - 5.2 < 0
    dbx, dby = np.abs(px - 11.0) - 4.0, np.abs(py + 6.0) - 3.0
    in_b = (dbx < 0) & (dby < 0)
    assert in_a.sum() > 50 and in_b.sum() > 30 and not (in_a & in_b).any()
    u, v = o.get("U0"), o.get("V0")
    want_u = np.full((H, W + 1), 123.0)
    want_v = np.full((H + 1, W), 123.0)
    for mask, vel in ((in_a, (3.5, -1.25)), (in_b, (-2.0, 6.0))):
        yy, xx = np.nonzero(mask)
        want_u[yy, xx] = vel[0]
        want_u[yy, xx + 1] = vel[0]
        want_v[yy, xx] = vel[1]
        want_v[yy + 1, xx] = vel[1]
    np.testing.assert_array_equal(u, want_u.astype(np.float32))
    np.testing.assert_array_equal(v, want_v.astype(np.float32))


def test_overlapping_sources_add_velocities_and_min_levels():
    W, H = 48, 40
    a = scenes.source((-3.0, 0.0), "circle", radius=6.0, velocity=(1.0, 2.0))
    b = scenes.source((3.0, 0.0), "circle", radius=6.0, velocity=(10.0, -4.0))
    o = _source_instance(W, H, [a, b], ls_value=20.0)
    px, py = _cell_world(W, H)
    da, db = np.hypot(px + 3.0, py) - 6.0, np.hypot(px - 3.0, py) - 6.0
    np.testing.assert_allclose(o.get("LS_AIR1"), np.minimum(20.0, np.minimum(da, db)), atol=2e-5)
    u = o.get("U0")
    # a cell deep in the overlap and its neighbours on both sides: 1 + 10
    both = (da < -1.5) & (db < -1.5)
    yy, xx = np.nonzero(both)
    assert len(yy) > 4
    assert (u[yy, xx] == 11.0).all() and (u[yy, xx + 1] == 11.0).all()


def test_aabb_source_level_closed_form():
    W, H = 48, 40
    c, hs = (2.0, -1.0), (9.0, 5.0)
    o = _source_instance(W, H, [scenes.source(c, "aabb", half_size=hs)], ls_value=1e6)
    px, py = _cell_world(W, H)
    dx, dy = np.abs(px - c[0]) - hs[0], np.abs(py - c[1]) - hs[1]
    inside = np.maximum(dx, dy)
    outside = np.hypot(np.maximum(dx, 0), np.maximum(dy, 0))
    want = np.where((dx < 0) & (dy < 0), inside, np.where(dx < 0, dy, np.where(dy < 0, dx, outside)))
    np.testing.assert_allclose(o.get("LS_AIR1"), want, atol=2e-5)


# ---------------------------------------------------------------------------
# sample_forces.wgsl + accumulate_forces.wgsl (+ fluid_to_solid.rs:200-204)
# ---------------------------------------------------------------------------
RHO, G, DXW = 99.7, 9.8, 0.1


def _hydrostatic_forces(ob, W=96, H=96, k=RHO * G):
    """Force / torque bins of one body in p = k * (y - 10) (grid y down)."""
    o = fresh(W, H)
    o.set_obstacles([ob])
    o.run_stage("UPDATE_SOLID", DT)
    o.run_stage("AREA_FRACTION", DT)
    ys = np.mgrid[0:H, 0:W][0].astype(np.float64)
    o.set("P0", (k * (ys - 10.0)).astype(np.float32))
    o.run_stage("FLUID_TO_SOLID", DT)
    af = o.get("AREA_FRACTION").astype(np.float64)
    # closed area the fractions describe: closed length of the horizontal
    # faces summed over the columns (each face counted once, as a +y face),
    # away from the wall ring
    closed = float(np.sum(1.0 - af[2:-2, 2:-2, 3]))
    band = int(((o.get("SOLID_ID") == ob.entity_id) & ~(af == 0).all(-1)).sum())
    return o.get("SOLID_FORCES")[0, ob.entity_id], o.get("SOLID_BINS")[:, ob.entity_id], closed, band


def test_fluid_to_solid_submerged_circle_buoyancy():
    R = 11.0
    ob = scenes.obstacle(0, 0, [R], 0.0, (0.0, -4.0))
    f, bins, closed, band = _hydrostatic_forces(ob)
    k = RHO * G
    # summation by parts: sum_c p_c (f_{+y} - f_{-y}) = k * closed area
    want_y = -k * closed * DXW
    trunc = 0.1 * band  # i32(v * 10) truncates: < 0.1 per cell and bin
    # the sum runs over the cells tagged with the body's id (centre within
    # 0.5 of the shape); a few corner cells with a clipped face are outside
    # it, which costs well under 1 %
    assert abs(f[1] - want_y) <= trunc + 0.006 * abs(want_y)
    assert abs(closed - math.pi * R * R) < 0.01 * math.pi * R * R
    assert abs(f[0]) <= trunc
    assert abs(f[2]) <= trunc  # centred: no torque
    # bins hold the same numbers x 10 as integers (fixed-point scale 0.1)
    assert bins[1] == int(round(f[1] * 10.0))
    assert f[1] < 0  # grid y points down: buoyancy is negative


def test_fluid_to_solid_off_centre_capsule_torque():
    """A capsule whose record centre (the transform's position) sits 8 cells
    left of its area centroid: torque = r.x f.y summed = -k A dx * 8."""
    r, a, b = 4.0, (2.0, 0.0), (14.0, 0.0)
    # entity_id is the record's index in the list (obstacle.rs:133-137):
    # sample_forces.wgsl reads obstacles[solid_id]
    ob = scenes.obstacle(0, 2, [a[0], a[1], b[0], b[1], r], 0.0, (-10.0, -2.0))
    f, bins, closed, band = _hydrostatic_forces(ob)
    k = RHO * G
    area = math.pi * r * r + 2 * r * (b[0] - a[0])
    assert abs(closed - area) < 0.02 * area
    want_t = -k * closed * DXW * 8.0
    trunc = 0.1 * band
    assert abs(f[2] - want_t) <= trunc + 0.01 * abs(want_t)
    assert abs(f[1] + k * closed * DXW) <= trunc + 0.006 * k * closed * DXW


# ---------------------------------------------------------------------------
# rigid-body integrator (no reference counterpart)
# ---------------------------------------------------------------------------
def _tank_with_body(density_ratio, W=128, H=128):
    # sides are multiples of 64: the x-/y-major face kernels cover 64 * floor(n / 64)
    # rows / columns (pipeline.rs:180-186), so other sizes leave faces without gravity
    proj = scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-7)
    s = scenes.settings(W, H, proj=proj, rein=scenes.reinit("fast_iterative", 10))
    o = OracleInstance(s)
    o.set_sources([scenes.source((0.0, -20.0), "aabb", half_size=(80.0, 60.0))], True)  # water up to world y 40
    # R = 11 on a cell centre: the cells the force band leaves out (centre
    # 0.5 - 0.71 from the circle, with a clipped corner) carry 0.3 % of it
    R = 11.0
    ob = scenes.obstacle(0, 0, [R], 0.0, (0.0, -12.0))
    o.set_obstacles([ob])
    o.set_bodies(scenes.bodies_for([ob], [density_ratio * RHO]))
    o.step(DT)  # initialisation
    return o, R


@pytest.mark.parametrize("ratio", [0.5, 2.0])
def test_body_first_step_buoyancy_ratio(ratio):
    o, R = _tank_with_body(ratio)
    o.step(DT)
    vy = o.get_obstacles(1)[0].linear_velocity[1]
    rb = ratio * RHO
    want = DT * (G / DXW) * (RHO - rb) / (rb + 4.0 * RHO)
    assert vy * want > 0  # sinks when heavier, rises when lighter
    assert abs(vy - want) < 0.05 * abs(want)


def test_neutrally_buoyant_body_stays_at_rest():
    heavy, _ = _tank_with_body(2.0)
    heavy.step(DT)
    v_heavy = abs(heavy.get_obstacles(1)[0].linear_velocity[1])
    o, _ = _tank_with_body(1.0)
    y0 = o.get_obstacles(1)[0].transform[13]
    for _ in range(3):
        o.step(DT)
    ob = o.get_obstacles(1)[0]
    assert abs(ob.linear_velocity[1]) < 0.05 * v_heavy
    assert abs(ob.transform[13] - y0) < 0.05 * v_heavy * 3 * DT + 1e-4
