"""CG reductions folded once per instance (k_prefold, k_cg_check) against
the fold every consuming block does for itself.

Instances whose grids have more than 2048 CG tiles take the one-block fold;
EFL_INLINE_FOLD_MAX=0 at instance creation forces it on small grids.  Both
use the same fixed-order fold, so a forced instance must step bit for bit as
an unforced one: every field and the solver statistics, for MGPCG on a batch
of instances of different sizes (tiled, bottom-only and coarse-only
preconditioners) and for the flexible CG of the learned preconditioner.
"""
import pytest

from bevy_eulerian_fluid_amd import native, scenes
from tests.parity_util import small_scene
from tests.test_gpu_graphs import _same, _setup

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def _say(msg):
    print(msg, flush=True)


def _make(ctx, sc, monkeypatch, forced):
    if forced:
        monkeypatch.setenv("EFL_INLINE_FOLD_MAX", "0")
    try:
        i = _setup(ctx, sc)
    finally:
        monkeypatch.delenv("EFL_INLINE_FOLD_MAX", raising=False)
    return i


def test_prefold_matches_inline_fold_mgpcg(monkeypatch):
    _say("one-block folds vs inline folds: MGPCG batch")
    p = scenes.projection("mgpcg", max_cycles=60, rel_tol=1e-5)
    scs = [scenes.batch_instance(scenes.SplitMix64(scenes.SEED_BASE + 3, k), proj=p) for k in range(3)]
    scs += [small_scene(200, 136, proj=p), small_scene(32, 24, proj=p)]
    ca, cb = native.Context(0), native.Context(0)
    try:
        ia, ib = [], []
        for sc in scs:
            for ctx, lst, forced in ((ca, ia, False), (cb, ib, True)):
                i = _make(ctx, sc, monkeypatch, forced)
                i.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"])
                                if "densities" in sc else None)
                lst.append(i)
        for k in range(5):
            ca.step(ia, DT)
            cb.step(ib, DT)
            for a, b in zip(ia, ib):
                _same(a, b, f"step {k} ({a.W}x{a.H})")
    finally:
        ca.close()
        cb.close()


def test_prefold_matches_inline_fold_learned(monkeypatch):
    _say("one-block folds vs inline folds: learned preconditioner 128x128")
    ctx = native.Context(0)
    try:
        sc = small_scene(128, 128, proj=scenes.projection("mgpcg", max_cycles=100, rel_tol=1e-5))
        a = _make(ctx, sc, monkeypatch, False)
        b = _make(ctx, sc, monkeypatch, True)
        ctx.set_obstacles(sc["obstacles"])
        net = native.Network(ctx)
        net.init_params(13)
        for i in (a, b):
            i.set_learned(net)
            i.set_projection(scenes.projection("learned", max_cycles=100, rel_tol=1e-5))
        for k in range(3):
            a.step(DT)
            b.step(DT)
            _same(a, b, f"step {k}")
            assert a.learned_iterations() == b.learned_iterations(), k
        net.close()
        a.close()
        b.close()
    finally:
        ctx.close()
