"""efl_autotune: it times the instance's own solve per candidate, leaves the
simulation state untouched, keeps the fastest candidate that meets the same
tolerance (never slower than the settings it started from, which are one of
the candidates), and the tuned instance keeps converging."""
import numpy as np
import pytest

from bevy_eulerian_fluid_amd import abi, native, scenes

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0
STATE = ("U0", "V0", "P0", "LS_AIR0", "LS_AIR1")  # derived fields (solids, divergence) are rebuilt by the next step


def _say(msg):
    print(msg, flush=True)


def _tank(ctx, size, proj):
    sc = scenes.coupled_tank(size, n_each=2, proj=proj)
    i = ctx.instance(sc["settings"])
    i.set_sources(sc["init_sources"], True)
    i.set_obstacles(sc["obstacles"], scenes.bodies_for(sc["obstacles"], sc["densities"]))
    for _ in range(6):
        i.step(DT)
    return i


def test_autotune_mgpcg():
    _say("autotune: MGPCG, coupled tank 256x256")
    ctx = native.Context(0)
    try:
        i = _tank(ctx, 256, scenes.projection("mgpcg", max_cycles=100, rel_tol=1e-5))
        before = {f: i.get(f).copy() for f in STATE}
        r = i.autotune(DT, apply=False)
        _say(f"  report only: {r}")
        for f in STATE:
            assert np.array_equal(i.get(f).view(np.uint32), before[f].view(np.uint32)), f
        assert r["best_ms"] <= r["default_ms"]
        assert r["trials"] >= 2 * 10
        assert r["bottom_cells"] in (1024, 4096)
        assert r["tile_h"] in (0, 8, 16)
        r = i.autotune(DT, apply=True)
        _say(f"  applied: {r}")
        for _ in range(3):
            i.step(DT)
            st = i.stats()
            assert st.converged == 1 and st.rel_residual <= 1e-5
        i.close()
    finally:
        ctx.close()


def test_autotune_reports_what_it_applies():
    """The reported tuning, reapplied with efl_apply_tuning to a fresh
    instance of the same scene, reproduces the tuned instance's steps bit for
    bit (same sweeps, bottom size and tile height), and both tile heights
    give the same fields as well (the smoother's result does not depend on
    its tiling)."""
    _say("autotune: reapply the reported tuning, 256x256, and compare tile heights")
    ctx = native.Context(0)
    try:
        proj = scenes.projection("mgpcg", max_cycles=100, rel_tol=1e-5)
        a, b, c = _tank(ctx, 256, proj), _tank(ctx, 256, proj), _tank(ctx, 256, proj)
        r = a.autotune(DT, apply=True)
        _say(f"  tuned: {r}")
        keep = {k: r[k] for k in ("pre", "post", "coarsest", "bottom_cells", "nn_strips", "tile_h")}
        b.apply_tuning(**keep)
        c.apply_tuning(**dict(keep, tile_h=8 if keep["tile_h"] == 16 else 16))  # another tile height
        for _ in range(4):
            for i in (a, b, c):
                i.step(DT)
        for f in STATE:
            ga = a.get(f)
            assert np.array_equal(ga.view(np.uint32), b.get(f).view(np.uint32)), f
            assert np.array_equal(ga.view(np.uint32), c.get(f).view(np.uint32)), f
        for i in (a, b, c):
            i.close()
    finally:
        ctx.close()


def test_autotune_rejects_fixed_solves():
    ctx = native.Context(0)
    try:
        i = _tank(ctx, 128, scenes.projection("gauss_seidel", gs_iterations=20))
        with pytest.raises(native.EflError) as e:
            i.autotune(DT)
        assert e.value.code == abi.E_STATE
        i.close()
    finally:
        ctx.close()


def test_autotune_learned_strips():
    _say("autotune: learned projection, network strips, 128x128")
    from bevy_eulerian_fluid_amd import learned
    ctx = native.Context(0)
    try:
        i = _tank(ctx, 128, scenes.projection("mgpcg", max_cycles=100, rel_tol=1e-5))
        net = native.Network(ctx)
        w = learned.load_weights()
        if w is None:
            net.init_params(3)
        else:
            net.set_params(w)
        i.set_learned(net)
        i.set_projection(scenes.projection("learned", max_cycles=100, rel_tol=1e-5))
        i.step(DT)
        r = i.autotune(DT, apply=True)
        _say(f"  applied: {r}")
        assert 0 <= r["nn_strips"] < 16
        i.step(DT)
        st = i.stats()
        assert st.converged == 1 and st.rel_residual <= 1e-5
        net.close()
        i.close()
    finally:
        ctx.close()
