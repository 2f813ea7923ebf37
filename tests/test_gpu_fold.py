"""CG reductions folded once per instance (k_prefold, k_cg_check) against
the fold every consuming block does for itself.

Instances whose grids have more than 2048 CG tiles take the one-block fold;
EFL_INLINE_FOLD_MAX=0 at instance creation forces it on small grids.  Both
use the same fixed-order fold, so a forced instance must step bit for bit as
an unforced one: every field and the solver statistics, for MGPCG on a batch
of instances of different sizes (tiled, bottom-only and coarse-only
preconditioners) and for the flexible CG of the learned preconditioner.
A forced instance runs its CG iterations as k_pcg_apply + k_pcg_axpy, so
the MGPCG comparison runs both contexts that way (EFL_CG_FUSE=0); the fused
iteration (Chronopoulos-Gear form, other rounding) is held to it within the
solver tolerance here too.
"""
import numpy as np
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
    _say("one-block folds vs inline folds: MGPCG batch, unfused iterations")
    p = scenes.projection("mgpcg", max_cycles=60, rel_tol=1e-5)
    scs = [scenes.batch_instance(scenes.SplitMix64(scenes.SEED_BASE + 3, k), proj=p) for k in range(3)]
    scs += [small_scene(200, 136, proj=p), small_scene(32, 24, proj=p)]
    monkeypatch.setenv("EFL_CG_FUSE", "0")
    try:
        ca, cb = native.Context(0), native.Context(0)
    finally:
        monkeypatch.delenv("EFL_CG_FUSE", raising=False)
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


@pytest.mark.parametrize("size", [(96, 80), (200, 136), (1024, 1024)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_fused_cg_matches_unfused(monkeypatch, size):
    """The fused iteration (one launch: p = z + beta p, q = A p, x, r, with
    alpha = r.z / (z.Az - beta r.z / alpha_prev)) against k_pcg_apply +
    k_pcg_axpy (alpha = r.z / p.q) on the same states: the same iterates in
    exact arithmetic, so both converge, their iteration counts differ by at
    most one and the pressures agree within 1e-5 of max |p| (the MGPCG
    parity bound)."""
    _say(f"fused vs unfused CG iterations {size}")
    p = scenes.projection("mgpcg", max_cycles=100, rel_tol=1e-5)
    if size[0] >= 1024:
        sc = scenes.coupled_tank(size[0], proj=p, n_each=1)
    else:
        sc = small_scene(*size, proj=p)
    monkeypatch.setenv("EFL_CG_FUSE", "0")
    try:
        cu = native.Context(0)
    finally:
        monkeypatch.delenv("EFL_CG_FUSE", raising=False)
    cf = native.Context(0)
    try:
        f, u = _setup(cf, sc), _setup(cu, sc)
        for c in (cf, cu):
            c.set_obstacles(sc["obstacles"])
        f.step(DT)  # initialisation
        u.step(DT)
        for fld in ("U0", "V0", "P0", "LS_AIR0"):
            f.set(fld, u.get(fld))
        for k in range(4):
            f.step(DT)
            u.step(DT)
            sf, su = f.stats(), u.stats()
            assert sf.converged == 1 and su.converged == 1, k
            assert abs(int(sf.cycles) - int(su.cycles)) <= 1, (k, sf.cycles, su.cycles)
            pf, pu = f.get("P0"), u.get("P0")
            scale = np.abs(pu).max()
            assert np.abs(pf - pu).max() <= 1e-5 * scale, (k, np.abs(pf - pu).max(), scale)
            for fld in ("U0", "V0", "P0", "LS_AIR0"):  # the next step starts from the same state
                f.set(fld, u.get(fld))
    finally:
        cf.close()
        cu.close()
