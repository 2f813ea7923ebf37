"""Replayed steps (HIP graphs) against eager launches.

Two contexts run the same scenes side by side, one replaying its steps as
HIP graphs, one launching every kernel (the default); after every step -
with host reads and synchronisation in between - every field and the solver
statistics must agree bit for bit.  Covers the MG-preconditioned CG (a
graph for the CG iteration replayed a data-dependent number of times), a
fixed Gauss-Seidel solve, two-way coupled bodies, and a batch of instances
of different sizes.  Wherever an instance's first two levels are tiled
(128x96, 96x80 and 128x64 here) the CG iteration runs fused into the
preconditioner's level-0 launch, in the replayed graphs (an odd and an even
iteration, whose residual buffers differ) as in eager launches, so the
replays are held to the eager iterations bit for bit; smaller instances of
the batch run k_pcg_apply + k_pcg_axpy.
"""
import numpy as np
import pytest

from bevy_eulerian_fluid_amd import native, scenes
from tests.parity_util import ALL_FIELDS, small_scene

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def _say(msg):
    print(msg, flush=True)


def _setup(ctx, sc):
    inst = ctx.instance(sc["settings"])
    inst.set_sources(sc.get("init_sources", []), True)
    inst.set_sources(sc.get("sources", []), False)
    return inst


def _same(a, b, where):
    for f in ALL_FIELDS:
        x, y = a.get(f), b.get(f)
        if x.dtype.kind == "f":
            ok = np.array_equal(x.view(np.uint32), y.view(np.uint32)) or (
                np.array_equal(np.isnan(x), np.isnan(y)) and np.array_equal(x[~np.isnan(x)], y[~np.isnan(y)]))
        else:
            ok = np.array_equal(x, y)
        assert ok, f"{where}: field {f} differs"
    sa, sb = a.stats(), b.stats()
    assert sa.cycles == sb.cycles and sa.converged == sb.converged, where
    assert sa.rel_residual == sb.rel_residual, where


@pytest.fixture()
def two_ctx():
    g = native.Context(0)
    g.set_graphs(True)
    e = native.Context(0)
    e.set_graphs(False)
    yield g, e
    g.close()
    e.close()


@pytest.mark.parametrize("proj", ["mgpcg", "gauss_seidel"])
def test_graph_replay_matches_eager(two_ctx, proj):
    _say(f"graphs vs eager {proj} 128x96")
    cg, ce = two_ctx
    p = scenes.projection(proj, gs_iterations=20, max_cycles=60, rel_tol=1e-6 if proj == "mgpcg" else 0.0)
    sc = small_scene(128, 96, proj=p, forces=False)
    ig, ie = _setup(cg, sc), _setup(ce, sc)
    for c in (cg, ce):
        c.set_obstacles(sc["obstacles"])
    for k in range(8):
        ig.step(DT)
        ie.step(DT)
        _same(ig, ie, f"step {k}")


def test_graph_replay_two_way_bodies(two_ctx):
    _say("graphs vs eager: coupled tank 128, dynamic bodies")
    cg, ce = two_ctx
    sc = scenes.coupled_tank(128, proj=scenes.projection("mgpcg", max_cycles=100, rel_tol=1e-5), n_each=1)
    bodies = scenes.bodies_for(sc["obstacles"], sc["densities"])
    ig, ie = _setup(cg, sc), _setup(ce, sc)
    for c in (cg, ce):
        c.set_obstacles(sc["obstacles"])
        c.set_bodies(bodies)
    n = len(sc["obstacles"])
    for k in range(10):
        ig.step(DT)
        ie.step(DT)
        _same(ig, ie, f"step {k}")
        og, oe = cg.get_obstacles(n), ce.get_obstacles(n)
        for a, b in zip(og, oe):
            assert list(a.transform) == list(b.transform), f"step {k}: body transform"


def test_graph_replay_batch(two_ctx):
    _say("graphs vs eager: batch of 3 instances")
    cg, ce = two_ctx
    p = scenes.projection("mgpcg", max_cycles=60, rel_tol=1e-5)
    scs = [small_scene(w, h, proj=p, forces=False, bodies=False) for (w, h) in [(64, 64), (96, 80), (128, 64)]]
    ig = [_setup(cg, sc) for sc in scs]
    ie = [_setup(ce, sc) for sc in scs]
    for k in range(6):
        cg.step(ig, DT)
        ce.step(ie, DT)
        for a, b in zip(ig, ie):
            _same(a, b, f"step {k}")


def test_batched_step_matches_single_steps():
    """A batch of instances of different sizes, each with its own two-way
    coupled bodies, stepped together in one efl_step (shared launches) gives
    bit for bit what stepping each instance alone gives."""
    _say("batched vs single instances")
    scs = [scenes.batch_instance(scenes.SplitMix64(scenes.SEED_BASE + 3, k),
                                 proj=scenes.projection("mgpcg", max_cycles=60, rel_tol=1e-5)) for k in range(4)]
    cb = native.Context(0)
    cs = [native.Context(0) for _ in scs]
    try:
        ib, isg = [], []
        for sc, c in zip(scs, cs):
            for ctx, lst in ((cb, ib), (c, isg)):
                i = _setup(ctx, sc)
                i.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"]))
                lst.append(i)
        for k in range(5):
            cb.step(ib, DT)
            for i in isg:
                i.step(DT)
            for a, b in zip(ib, isg):
                _same(a, b, f"step {k}")
    finally:
        cb.close()
        for c in cs:
            c.close()


def test_mixed_batch_matches_single_steps():
    """A batch holding a 1024^2 instance (lazy stencils on level 0) beside four
    smaller ones whose V-cycles bottom out at different levels and whose
    solves end at different iterations, stepped in one efl_step: the level-0
    launches split into a lazy and a plain launch, every bottom solve runs in
    one launch, and each instance is queued only its own iterations.  Every
    field bit for bit what stepping each instance alone gives."""
    _say("mixed batch (1024^2 + 4 small) vs single instances")
    p = scenes.projection("mgpcg", max_cycles=60, rel_tol=1e-5)
    scs = [scenes.batch_instance(scenes.SplitMix64(scenes.SEED_BASE + 3, 7), proj=p, size=1024)]
    scs += [scenes.batch_instance(scenes.SplitMix64(scenes.SEED_BASE + 3, k), proj=p) for k in range(4)]
    cb = native.Context(0)
    cs = [native.Context(0) for _ in scs]
    try:
        ib, isg = [], []
        for sc, c in zip(scs, cs):
            for ctx, lst in ((cb, ib), (c, isg)):
                i = _setup(ctx, sc)
                i.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"]))
                lst.append(i)
        for k in range(4):
            cb.step(ib, DT)
            for i in isg:
                i.step(DT)
            its = [i.stats().cycles for i in isg]
            for a, b in zip(ib, isg):
                _same(a, b, f"step {k}")
                assert a.stats().cycles == b.stats().cycles
        _say(f"  last step's iterations per instance: {its}")
    finally:
        cb.close()
        for c in cs:
            c.close()


@pytest.mark.parametrize("proj", ["mgpcg", "multigrid"])
@pytest.mark.parametrize("size", [(200, 136), (1024, 1024)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_up_pairs_match_separate_launches(monkeypatch, size, proj):
    """The up pass in level pairs (k_mg_up2: level l + 1 smoothed inside
    level l's launch) against one launch per level (EFL_UP_MERGE=0): every
    field and the solver statistics bit for bit over a few steps."""
    _say(f"paired vs separate up launches {proj} {size}")
    p = scenes.projection(proj, max_cycles=60 if proj == "mgpcg" else 2, rel_tol=1e-5 if proj == "mgpcg" else 0.0)
    sc = small_scene(*size, proj=p, forces=False)
    monkeypatch.setenv("EFL_UP_MERGE", "0")
    try:
        cs = native.Context(0)
    finally:
        monkeypatch.delenv("EFL_UP_MERGE", raising=False)
    cp = native.Context(0)
    try:
        a, b = _setup(cp, sc), _setup(cs, sc)
        for c in (cp, cs):
            c.set_obstacles(sc["obstacles"])
        for k in range(4):
            a.step(DT)
            b.step(DT)
            _same(a, b, f"step {k}")
    finally:
        cp.close()
        cs.close()
