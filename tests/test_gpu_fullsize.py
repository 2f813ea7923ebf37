"""The benchmark workloads at their full sizes, checked through properties
that do not depend on the size (the oracle is too slow to replay them step
by step): every pressure solve reaches its tolerance, the normalised
divergence after the projection stays small, the fluid volume (cells with
ls_air < 0, as the reference's diagnostics count it) drifts little, and
every field stays finite.  Smaller sizes of the same scenes are compared
with the oracle bit for bit in the other GPU tests."""
import numpy as np
import pytest

from bevy_eulerian_fluid_amd import native, scenes

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0
FIELDS = ("U0", "V0", "P0", "LS_AIR0", "LS_SOLID")


def _say(msg):
    print(msg, flush=True)


def _finite(inst, where):
    for f in FIELDS:
        a = inst.get(f)
        assert np.isfinite(a).all(), f"{where}: {f} not finite"


def test_coupled_tank_1024_properties():
    _say("coupled tank 1024^2, 16 two-way coupled bodies, MGPCG to 1e-5: 40 steps")
    sc = scenes.coupled_tank(1024, proj=scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-5, check_every=1))
    ctx = native.Context(0)
    inst = ctx.instance(sc["settings"])
    try:
        inst.set_sources(sc["init_sources"], True)
        ctx.set_obstacles(sc["obstacles"])
        ctx.set_bodies(scenes.bodies_for(sc["obstacles"], sc["densities"]))
        inst.set_diagnostics(True)
        inst.step(DT)
        v0 = inst.fluid_diagnostics()[0]
        assert v0 > 0.4 * 1024 * 1024
        for k in range(1, 40):
            inst.step(DT)
            st = inst.stats()
            assert st.converged == 1 and 0.0 <= st.rel_residual <= 1e-5, (k, st.cycles, st.rel_residual)
            assert st.max_norm_divergence < 1e-3, (k, st.max_norm_divergence)
        vol, vmin, vmax = inst.fluid_diagnostics()
        assert abs(vol - v0) / v0 < 0.05, (v0, vol)
        # as in the reference, each of the 512 lanes of the second pass adds
        # its 8 block values: vmax is a sum of block maxima, and vmin a sum of
        # block minima in which every all-air block counts 1e30
        assert 0.0 < vmax < 1e4 and vmin >= 0.0
        _finite(inst, "tank")
        for o in ctx.get_obstacles(len(sc["obstacles"])):
            assert np.isfinite(np.array(list(o.transform))).all()
    finally:
        inst.close()
        ctx.close()


def test_combined_4096_batch_properties():
    _say("4096^2 droplets + pool + 24 bodies + 64 forces, batched with 8 x 512^2: 4 steps")
    sc = scenes.combined(4096)
    ctx = native.Context(0)
    insts = []
    try:
        big = ctx.instance(sc["settings"])
        insts.append(big)
        big.init_levelset(sc["levelset"])
        ctx.set_obstacles(sc["obstacles"])
        ctx.set_bodies(scenes.bodies_for(sc["obstacles"], sc["densities"]))
        for k in range(8):
            s2 = scenes.coupled_tank(512, seed=scenes.SEED_BASE + 50 + k, n_each=1,
                                     proj=scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-5))
            i = ctx.instance(s2["settings"])
            i.set_sources(s2["init_sources"], True)
            i.set_obstacles(s2["obstacles"], scenes.bodies_for(s2["obstacles"], s2["densities"]))
            insts.append(i)
        big.set_forces(sc["forces"])
        ctx.step(insts, DT)  # the first step of an instance initialises it (no projection)
        for k in range(4):
            ctx.step(insts, DT)
            for j, i in enumerate(insts):
                st = i.stats()
                assert st.converged == 1, (k, j, st.cycles, st.rel_residual)
        vol = big.fluid_diagnostics()[0]
        assert 0.2 * 4096 * 4096 * 0.9 < vol < 0.6 * 4096 * 4096
        for j, i in enumerate(insts):
            _finite(i, f"instance {j}")
    finally:
        for i in insts:
            i.close()
        ctx.close()
