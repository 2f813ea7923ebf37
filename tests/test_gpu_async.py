"""Asynchronous MGPCG steps (efl_set_async).

An asynchronous context queues each solve's iteration budget (the largest
recent count plus a margin) and never waits for the step's own solve; it keeps
at most two steps in flight per instance; statistics resolve at efl_sync /
efl_get_stats.  Iterations queued after
the device-side stopping test fired return at once, so with a budget that
covers every solve the fields are those of a synchronous context bit for bit.
A budget that does not cover the solve stops it there, unconverged, and says
so in the statistics.
"""
import numpy as np
import pytest

from bevy_eulerian_fluid_amd import native, scenes

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0
FIELDS = ("U0", "V0", "P0", "LS_AIR0", "LS_SOLID")


def _say(msg):
    print(msg, flush=True)


def _tank(ctx, size=256, rel_tol=1e-5):
    sc = scenes.coupled_tank(size, proj=scenes.projection("mgpcg", max_cycles=200, rel_tol=rel_tol), n_each=2)
    i = ctx.instance(sc["settings"])
    i.set_sources(sc["init_sources"], True)
    ctx.set_obstacles(sc["obstacles"])
    ctx.set_bodies(scenes.bodies_for(sc["obstacles"], sc["densities"]))
    return i


def test_async_steps_match_sync_and_do_not_wait():
    _say("async MGPCG: 256^2 coupled tank, 12 steps, margin 16, against a synchronous context")
    a_ctx, b_ctx = native.Context(0), native.Context(0)
    try:
        a, b = _tank(a_ctx), _tank(b_ctx)
        b_ctx.set_async(16)
        a.step(DT)
        b.step(DT)
        b_ctx.sync()
        cyc_a = []
        for k in range(12):
            a.step(DT)
            cyc_a.append(int(a.stats().cycles))
            w0 = b_ctx.host_waits()
            b.step(DT)  # the device is idle: nothing in flight, so no wait at all
            assert b_ctx.host_waits() == w0, "an asynchronous step waited for the device"
            b_ctx.sync()
        st = b.stats()
        assert st.converged == 1 and st.rel_residual <= 1e-5, (st.cycles, st.rel_residual)
        assert st.unconverged_solves == 0 and a.stats().unconverged_solves == 0
        assert int(st.cycles) == cyc_a[-1]
        for f in FIELDS:
            ga, gb = a.get(f), b.get(f)
            assert np.array_equal(ga, gb), f
        assert a_ctx.host_waits() > 0  # the synchronous context does read back
        for x, y in zip(a_ctx.get_obstacles(8), b_ctx.get_obstacles(8)):
            assert list(x.transform) == list(y.transform)
        _say(f"  iterations per step (sync): {cyc_a}")
        # free-running: at most one wait per step (for the step before last),
        # none while fewer than two steps are in flight
        w0 = b_ctx.host_waits()
        for k in range(20):
            b.step(DT)
        assert b_ctx.host_waits() - w0 <= 18
        b_ctx.sync()
        assert b.stats().converged == 1
    finally:
        a_ctx.close()
        b_ctx.close()


def test_async_budget_too_small_reports_unconverged():
    _say("async MGPCG with margin 1: a budget below the solve stops unconverged")
    ctx = native.Context(0)
    try:
        i = _tank(ctx, rel_tol=1e-7)
        ctx.set_async(1)
        i.step(DT)
        budgets = []
        for _ in range(3):
            i.step(DT)
            ctx.sync()  # resolves the step's statistics
            st = i.stats()
            budgets.append((int(st.cycles), int(st.converged), float(st.rel_residual)))
            assert (st.converged == 1) == (st.rel_residual <= 1e-7), budgets
        # the first budget is 4 + 1 iterations; 1e-7 needs more, and each
        # following budget is the previous count plus the margin
        assert budgets[0][0] == 5 and budgets[0][1] == 0, budgets
        # counted on the device: one per step that ended unconverged
        assert i.stats().unconverged_solves == sum(1 - b[1] for b in budgets), budgets
        assert budgets[1][0] <= 6 and budgets[2][0] <= budgets[1][0] + 1, budgets
        for f in FIELDS:
            assert np.isfinite(i.get(f)).all(), f
    finally:
        ctx.close()


def test_async_batch_matches_sync():
    """A batch of four instances whose solves end at different iterations,
    asynchronous with a margin that covers every solve (each instance queued
    its own budget) against the same batch stepped synchronously: every
    field bit for bit, every solve converged."""
    _say("async MGPCG: batch of 4 instances, margin 16, against a synchronous batch")
    p = scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-5)
    scs = [scenes.batch_instance(scenes.SplitMix64(scenes.SEED_BASE + 3, k), proj=p) for k in range(4)]
    a_ctx, b_ctx = native.Context(0), native.Context(0)
    try:
        ia, ib = [], []
        for ctx, lst in ((a_ctx, ia), (b_ctx, ib)):
            for sc in scs:
                i = ctx.instance(sc["settings"])
                i.set_sources(sc["init_sources"], True)
                i.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"]))
                lst.append(i)
        b_ctx.set_async(16)
        for k in range(8):
            a_ctx.step(ia, DT)
            b_ctx.step(ib, DT)
        b_ctx.sync()
        its = [int(i.stats().cycles) for i in ia]
        for a, b in zip(ia, ib):
            sa, sb = a.stats(), b.stats()
            assert sb.converged == 1 and sb.unconverged_solves == 0, (sb.cycles, sb.rel_residual)
            assert int(sa.cycles) == int(sb.cycles)
            for f in FIELDS:
                assert np.array_equal(a.get(f), b.get(f)), f
        _say(f"  last step's iterations per instance: {its}")
    finally:
        a_ctx.close()
        b_ctx.close()
