"""GPU parity: libefl.so (HIP, gfx950) against the CPU oracle.

Per stage: the oracle runs a few whole steps and then the stages that
precede the one under test, its complete state is copied into the GPU
instance, and both run the stage.  Every field must then agree bit for bit,
except where the stage calls exp() (local forces: ocml vs glibc, 2e-6
relative) or reduces in a different order (MGPCG: 1e-5 relative on p0).
"""
import numpy as np
import pytest

from bevy_eulerian_fluid_amd import abi, native, scenes
from oracle.oracle import OracleInstance
from tests.parity_util import ALL_FIELDS, compare_all, copy_state, make_pair, small_scene

NO_R0 = [f for f in ALL_FIELDS if f != "MG_R0"]

pytestmark = pytest.mark.gpu

STEP_ORDER = ["UPDATE_SOLID", "AREA_FRACTION", "ADVECT_VELOCITY", "APPLY_FORCES", "DIVERGENCE",
              "PROJECTION", "SOLVE_VELOCITY", "EXTRAPOLATE_VELOCITY", "ADVECT_LEVELSET", "SOURCE",
              "REINITIALIZE", "LEVELSET_GRADIENT", "FLUID_TO_SOLID"]
SIZES = [(64, 64), (96, 80), (100, 70), (200, 136)]
DT = 1.0 / 60.0


def _say(msg):
    print(msg, flush=True)


def _warm_oracle(o, sc, steps=3):
    o.set_forces(sc["forces"])
    o.step(DT)  # initialisation
    for _ in range(steps):
        o.set_forces(sc["forces"])
        o.step(DT)


@pytest.mark.parametrize("size", SIZES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("stage", STEP_ORDER)
def test_stage_parity(gpu_ctx, size, stage):
    _stage_case(gpu_ctx, size, stage)


# Stages whose kernels stage tiles plus halos in LDS and write their fields
# in place: at these sizes a launch has more tiles than run at once, so late
# tiles read halos that early tiles may already have written.
@pytest.mark.parametrize("size", [(1024, 1024), (1000, 700)], ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("stage", ["EXTRAPOLATE_VELOCITY", "REINITIALIZE"])
def test_stage_parity_multiwave(gpu_ctx, size, stage):
    _stage_case(gpu_ctx, size, stage)


@pytest.mark.parametrize("size", [(64, 64), (200, 136), (1000, 700)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_fluid_to_solid_parity_submerged(gpu_ctx, size):
    """Every body (circle, rectangle, capsule, triangle) under water, so that
    each one's bins hold forces and torques (about the triangle's centroid,
    the others' translations); 1000 x 700 leaves partial tiles on two edges."""
    bins = _stage_case(gpu_ctx, size, "FLUID_TO_SOLID", small_scene(*size, submerged=True))
    assert (bins[:, :4] != 0).all(), bins[:, :4]


def _stage_case(gpu_ctx, size, stage, sc=None):
    _say(f"stage parity {stage} {size}")
    sc = sc or small_scene(*size)
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    try:
        _warm_oracle(o, sc)
        o.set_forces(sc["forces"])
        for st in STEP_ORDER[:STEP_ORDER.index(stage)]:
            o.run_stage(st, DT)
        copy_state(o, g)
        g.set_forces(sc["forces"])
        o.run_stage(stage, DT)
        g.run_stage(stage, DT)
        tol = {}
        if stage == "APPLY_FORCES":
            tol = {"U1": (2e-6, 1e-6), "V1": (2e-6, 1e-6)}
        errs = compare_all(g, o, tol=tol)
        assert not errs, "\n".join(errs)
        return g.get("SOLID_BINS")
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("stage", ["INIT_GRID_CENTER", "INIT_GRID_EDGE", "SOURCE_ONESHOT"])
@pytest.mark.parametrize("size", SIZES + [(64, 128), (128, 64)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_init_stage_parity(gpu_ctx, size, stage):
    _say(f"init stage parity {stage} {size}")
    sc = small_scene(*size)
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    try:
        # start from non-zero garbage so that untouched cells are visible
        rng = np.random.default_rng(1)
        for f in ["U0", "U1", "V0", "V1", "LS_AIR0", "LS_AIR1", "GRAD_LS"]:
            a = rng.standard_normal(o.get(f).shape).astype(np.float32)
            o.set(f, a)
            g.set(f, a)
        o.run_stage(stage, DT)
        g.run_stage(stage, DT)
        errs = compare_all(g, o)
        assert not errs, "\n".join(errs)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("method", ["gauss_seidel", "jacobi", "multigrid", "gs_sweep", "mg_vcycle",
                                    "mg_residual", "mgpcg", "mg_smooth", "gs_sweep_split", "multigrid_3cycles"])
@pytest.mark.parametrize("size", SIZES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_projection_parity(gpu_ctx, size, method):
    _projection_case(gpu_ctx, size, method)


# Grids whose finest level needs more tiles than run at once: a tiled
# smoother launch then has blocks that start after others have finished, so
# any block reading a halo that another block of the same launch writes
# would show up here as a mismatch.
@pytest.mark.parametrize("method", ["gauss_seidel", "gs_sweep", "mg_vcycle", "mg_smooth", "multigrid_3cycles"])
@pytest.mark.parametrize("size", [(1024, 1024), (1000, 700)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_projection_parity_multiwave(gpu_ctx, size, method):
    _projection_case(gpu_ctx, size, method)


def _projection_case(gpu_ctx, size, method, tile_h=None, bottom_cells=1024):
    _say(f"projection parity {method} {size} tile_h {tile_h} bottom {bottom_cells}")
    stage = "PROJECTION"
    proj = scenes.projection("multigrid", max_cycles=1)
    if method in ("gauss_seidel", "jacobi"):
        proj = scenes.projection(method, gs_iterations=30)
    elif method == "mgpcg":
        proj = scenes.projection("mgpcg", max_cycles=50, rel_tol=1e-5)
    elif method == "gs_sweep":
        stage = "GS_SWEEP"
    elif method == "mg_vcycle":
        stage = "MG_VCYCLE"
    elif method == "mg_residual":
        stage = "MG_RESIDUAL"
    elif method == "mg_smooth":
        stage = "MG_SMOOTH"
    elif method == "gs_sweep_split":
        stage = "GS_SWEEP_SPLIT"
    elif method == "multigrid_3cycles":
        proj = scenes.projection("multigrid", pre=3, post=1, coarsest=4, max_cycles=3)
    sc = small_scene(*size, proj=proj)
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    if tile_h is not None or bottom_cells != 1024:  # else the defaults
        g.apply_tuning(pre=proj.mg_pre_iterations, post=proj.mg_post_iterations,
                       coarsest=proj.mg_coarsest_iterations, tile_h=tile_h or 0, bottom_cells=bottom_cells)
    try:
        _warm_oracle(o, sc)
        for st in STEP_ORDER[:STEP_ORDER.index("PROJECTION")]:
            o.run_stage(st, DT)
        copy_state(o, g)
        o.run_stage(stage, DT)
        g.run_stage(stage, DT)
        if method == "mgpcg":
            po, pg = o.get("P0"), g.get("P0")
            scale = np.abs(po).max()
            assert np.abs(pg - po).max() <= 1e-5 * scale, (np.abs(pg - po).max(), scale)
            so, sg = o.stats(), g.stats()
            assert sg.converged == 1 and so.converged == 1
            assert abs(int(sg.cycles) - int(so.cycles)) <= 1
        else:
            # the fused V-cycle never materialises the level-0 residual
            fields = ALL_FIELDS if stage == "MG_RESIDUAL" else NO_R0
            errs = compare_all(g, o, fields)
            assert not errs, "\n".join(errs)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("reinit", ["fast_iterative", "jump_flooding"])
@pytest.mark.parametrize("size", SIZES + [(128, 64)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_reinit_parity(gpu_ctx, size, reinit):
    _say(f"reinit parity {reinit} {size}")
    sc = small_scene(*size, rein=scenes.reinit(reinit, 10))
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    try:
        _warm_oracle(o, sc)
        for st in STEP_ORDER[:STEP_ORDER.index("REINITIALIZE")]:
            o.run_stage(st, DT)
        copy_state(o, g)
        o.run_stage("REINITIALIZE", DT)
        g.run_stage("REINITIALIZE", DT)
        errs = compare_all(g, o)
        assert not errs, "\n".join(errs)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("size", SIZES, ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("proj", ["gauss_seidel", "multigrid"])
def test_multistep_bitexact(gpu_ctx, size, proj):
    """Initialisation + 12 whole steps without local forces: bit-exact."""
    _say(f"multistep {proj} {size}")
    p = scenes.projection(proj, gs_iterations=20, max_cycles=1)
    sc = small_scene(*size, proj=p, forces=False)
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    try:
        for k in range(13):
            o.step(DT)
            g.step(DT)
            errs = compare_all(g, o, NO_R0)
            assert not errs, f"step {k}: " + "\n".join(errs)
    finally:
        g.close()
        o.close()


def test_bodies_two_way_parity(gpu_ctx):
    """Dynamic bodies integrated on the device follow the oracle's
    trajectories (transcendentals differ by ulps, so within tolerance)."""
    _say("bodies two-way 128x128")
    sc = scenes.coupled_tank(128, proj=scenes.projection("mgpcg", max_cycles=100, rel_tol=1e-6), n_each=1)
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    bodies = scenes.bodies_for(sc["obstacles"], sc["densities"])
    gpu_ctx.set_bodies(bodies)
    o.set_bodies(bodies)
    try:
        for _ in range(16):
            g.step(DT)
            o.step(DT)
        n = len(sc["obstacles"])
        og, oo = gpu_ctx.get_obstacles(n), o.get_obstacles(n)
        for a, b in zip(og, oo):
            assert abs(a.transform[12] - b.transform[12]) < 1e-2
            assert abs(a.transform[13] - b.transform[13]) < 1e-2
            assert abs(a.linear_velocity[1] - b.linear_velocity[1]) < 1e-2 * (1 + abs(b.linear_velocity[1]))
        moved = max(abs(b.linear_velocity[1]) for b in oo)
        assert moved > 0.1
    finally:
        gpu_ctx.set_bodies([])
        gpu_ctx.set_obstacles([])
        g.close()
        o.close()


@pytest.mark.parametrize("size", [(64, 64), (200, 136)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_mgpcg_stops_on_non_finite_state(gpu_ctx, size):
    """A NaN in the pressure guess at an unknown makes r, and then p . q, not
    a number: the oracle's CG loop breaks before its first update, so the GPU
    solve must stop there too (0 iterations, not converged, pressure left as
    it was) instead of running max_cycles iterations."""
    _say(f"mgpcg non-finite state {size}")
    proj = scenes.projection("mgpcg", max_cycles=50, rel_tol=1e-5)
    sc = small_scene(*size, proj=proj)
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    try:
        _warm_oracle(o, sc)
        for st in STEP_ORDER[:STEP_ORDER.index("PROJECTION")]:
            o.run_stage(st, DT)
        ls, af = o.get("LS_AIR0"), o.get("AREA_FRACTION")
        open_fluid = (ls < -2.0) & (af.sum(axis=2) == 4.0)
        y, x = np.argwhere(open_fluid)[0]
        p = o.get("P0")
        p[y, x] = np.nan
        o.set("P0", p)
        copy_state(o, g)
        o.run_stage("PROJECTION", DT)
        g.run_stage("PROJECTION", DT)
        so, sg = o.stats(), g.stats()
        assert int(so.cycles) == 0 and so.converged == 0
        assert int(sg.cycles) == int(so.cycles) and sg.converged == so.converged
        assert np.isnan(sg.rel_residual)
        errs = compare_all(g, o, ["P0"])
        assert not errs, "\n".join(errs)
    finally:
        g.close()
        o.close()


@pytest.mark.parametrize("after", ["gauss_seidel", "multigrid", "mg_smooth", "gs_sweep"])
@pytest.mark.parametrize("size", [(128, 128), (200, 136)], ids=lambda s: f"{s[0]}x{s[1]}")
@pytest.mark.parametrize("cg_cycles", [1, 3])
def test_projection_after_mgpcg_switch(gpu_ctx, size, after, cg_cycles):
    """An MGPCG solve whose fused x / r update ran an odd number of times
    leaves the CG residual in its second buffer.  A later Gauss-Seidel or
    multigrid solve of the same instance must still smooth against the
    divergence: switching away from MGPCG is bit-exact with the oracle."""
    _say(f"projection after mgpcg ({cg_cycles} cycles) -> {after} {size}")
    cg = scenes.projection("mgpcg", max_cycles=cg_cycles, rel_tol=1e-9)
    sc = small_scene(*size, proj=cg)
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    try:
        _warm_oracle(o, sc)
        for st in STEP_ORDER[:STEP_ORDER.index("PROJECTION")]:
            o.run_stage(st, DT)
        copy_state(o, g)
        g.run_stage("PROJECTION", DT)  # max_cycles odd iterations, each carrying the fused update
        assert int(g.stats().cycles) == cg_cycles
        stage = {"mg_smooth": "MG_SMOOTH", "gs_sweep": "GS_SWEEP"}.get(after, "PROJECTION")
        proj = {"gauss_seidel": scenes.projection("gauss_seidel", gs_iterations=30),
                "multigrid": scenes.projection("multigrid", max_cycles=2)}.get(after)
        if proj is not None:
            g.set_projection(proj)
            o.set_projection(proj)
        copy_state(o, g)
        o.run_stage(stage, DT)
        g.run_stage(stage, DT)
        errs = compare_all(g, o, NO_R0)
        assert not errs, "\n".join(errs)
    finally:
        g.close()
        o.close()


def test_mgpcg_sealed_region_repair(gpu_ctx):
    """A fluid disc sealed by a solid ring with divergence that does not sum
    to zero: MGPCG cannot converge on it, so the library (as the oracle)
    removes the region's mean divergence and solves again.  The repaired
    divergence is bit-exact with the oracle's, the pressure within MGPCG's
    tolerance, and both solves converge."""
    from tests.sealed_scene import sealed_scene
    _say("mgpcg sealed region 64x64")
    o, _ = sealed_scene(OracleInstance)
    g = gpu_ctx.instance(o._settings)
    try:
        copy_state(o, g)
        o.run_stage("PROJECTION", DT)
        g.run_stage("PROJECTION", DT)
        so, sg = o.stats(), g.stats()
        assert so.converged == 1 and sg.converged == 1, (so.cycles, sg.cycles, sg.rel_residual)
        errs = compare_all(g, o, ["DIV"])
        assert not errs, "\n".join(errs)
        po, pg = o.get("P0"), g.get("P0")
        assert np.isfinite(pg).all()
        assert np.abs(pg - po).max() <= 1e-5 * np.abs(po).max()
    finally:
        g.close()
        o.close()


@pytest.fixture
def fused_ctx():
    """A context whose V-cycles run their coarse levels and bottom as one
    persistent launch (k_mg_coarse, EFL_COARSE_FUSE=1 at creation)."""
    import os
    from bevy_eulerian_fluid_amd import native
    old = os.environ.get("EFL_COARSE_FUSE")
    os.environ["EFL_COARSE_FUSE"] = "1"
    try:
        ctx = native.Context(0)
    finally:
        if old is None:
            del os.environ["EFL_COARSE_FUSE"]
        else:
            os.environ["EFL_COARSE_FUSE"] = old
    yield ctx
    ctx.close()


@pytest.mark.parametrize("method", ["mg_vcycle", "multigrid", "mgpcg"])
@pytest.mark.parametrize("size", [(100, 70), (200, 136), (1024, 1024), (1000, 700)],
                         ids=lambda s: f"{s[0]}x{s[1]}")
def test_projection_parity_fused_coarse_cycle(fused_ctx, size, method):
    """The persistent coarse-cycle launch does each phase's work with the same
    expressions as the separate launches: bit-exact with the oracle (MGPCG:
    its usual tolerance)."""
    _projection_case(fused_ctx, size, method)


@pytest.mark.parametrize("tile_h", [16, 8])
@pytest.mark.parametrize("method", ["gs_sweep", "mg_smooth", "mg_vcycle", "multigrid", "multigrid_3cycles", "mgpcg"])
@pytest.mark.parametrize("size", [(100, 70), (200, 136), (1000, 700), (1024, 1024)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_projection_parity_tile_h(gpu_ctx, size, method, tile_h):
    """Smoother launches of 64 x 16 or 64 x 8 cells per workgroup on every
    level (efl_apply_tuning tile_h; the default picks by level size): every
    cell's update reads the same values either way, so the result is
    bit-exact with the oracle (MGPCG: its usual tolerance; its level 0 keeps
    64 x 16)."""
    _projection_case(gpu_ctx, size, method, tile_h=tile_h)


@pytest.mark.parametrize("method", ["mg_vcycle", "multigrid", "multigrid_3cycles", "mgpcg"])
@pytest.mark.parametrize("size", [(100, 70), (200, 136), (1000, 700), (1024, 1024)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_projection_parity_bottom_4096(gpu_ctx, size, method):
    """The bottom solve from the first level of at most 4096 cells (64 x 64
    at 1024^2; 50 x 35 and 50 x 34 at the small sizes): two levels for the
    whole workgroup in register blocks, including partial blocks at odd
    sizes, then the wavefront-0 tail: bit-exact with the oracle (MGPCG:
    its usual tolerance)."""
    _projection_case(gpu_ctx, size, method, bottom_cells=4096)


def test_mixed_tile_heights_in_one_batch(gpu_ctx):
    """Two instances of one scene, tile heights 16 and 8, stepped together by
    one efl_step (the scheduler splits the smoother launches by tile height):
    both match the oracle bit for bit over 6 multigrid steps."""
    _say("mixed tile heights in one batched step, 200x136, 6 steps")
    p = scenes.projection("multigrid", max_cycles=1)
    sc = small_scene(200, 136, proj=p, forces=False)
    g, o = make_pair(gpu_ctx, sc, OracleInstance)
    h = gpu_ctx.instance(sc["settings"])
    h.set_sources(sc.get("init_sources", []), True)
    h.set_sources(sc.get("sources", []), False)
    g.apply_tuning(tile_h=16)
    h.apply_tuning(tile_h=8)
    try:
        for k in range(7):
            o.step(DT)
            gpu_ctx.step([g, h], DT)
            for inst, name in ((g, "tile_h 16"), (h, "tile_h 8")):
                errs = compare_all(inst, o, NO_R0)
                assert not errs, f"step {k} {name}: " + "\n".join(errs)
    finally:
        g.close()
        h.close()
        o.close()


def test_apply_tuning_rejects_out_of_range(gpu_ctx):
    sc = small_scene(96, 80, proj=scenes.projection("multigrid", max_cycles=1))
    g = gpu_ctx.instance(sc["settings"])
    try:
        for bad in ({"tile_h": 4}, {"tile_h": 32}, {"pre": 0}, {"post": 9}, {"bottom_cells": 2048}):
            with pytest.raises(native.EflError) as e:
                g.apply_tuning(**bad)
            assert e.value.code == abi.E_INVALID, bad
        for ok in (0, 8, 16):
            g.apply_tuning(tile_h=ok)
    finally:
        g.close()


@pytest.mark.parametrize("size", [(96, 80), (200, 136)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_multistep_bitexact_fused_coarse_cycle(fused_ctx, size):
    """Initialisation + 12 multigrid steps through the fused coarse cycle."""
    _say(f"multistep fused coarse cycle {size}")
    p = scenes.projection("multigrid", max_cycles=1)
    sc = small_scene(*size, proj=p, forces=False)
    g, o = make_pair(fused_ctx, sc, OracleInstance)
    try:
        for k in range(13):
            o.step(DT)
            g.step(DT)
            errs = compare_all(g, o, NO_R0)
            assert not errs, f"step {k}: " + "\n".join(errs)
    finally:
        g.close()
        o.close()
