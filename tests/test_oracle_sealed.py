"""The MGPCG solve on a system with a sealed, incompatible fluid region (no
GPU): the oracle makes the divergence compatible there (its mean removed)
and converges; everything outside the region keeps its divergence."""
import numpy as np

from oracle.oracle import OracleInstance
from tests.sealed_scene import DT, sealed_scene


def test_sealed_region_gets_compatible_divergence_and_converges():
    o, r = sealed_scene(OracleInstance)
    d0 = o.get("DIV").copy()
    af = o.get("AREA_FRACTION")
    ls = o.get("LS_AIR0")
    act = (ls < 0) & ~(af == 0).all(-1)
    inside = act & (r < 7.0)  # the disc the ring seals
    assert inside.sum() > 100 and abs(d0[inside].sum()) > 10.0
    o.run_stage("PROJECTION", DT)
    st = o.stats()
    assert st.converged == 1 and st.rel_residual <= 1e-5
    p = o.get("P0")
    assert np.isfinite(p).all()
    d1 = o.get("DIV")
    pocket = act & (r < 9.0)
    # the pocket's divergence now sums to zero (float32 rounding of the
    # per-cell subtraction aside) and differs from the old one by a constant
    assert abs(d1[pocket].astype(np.float64).sum()) < 1e-4 * np.abs(d0[pocket]).sum()
    shift = (d0[pocket] - d1[pocket]).astype(np.float64)
    np.testing.assert_allclose(shift, np.mean(d0[pocket].astype(np.float64)), rtol=1e-5, atol=1e-6)
    # outside it nothing changed
    assert np.array_equal(d1[~pocket], d0[~pocket])
    # the solution satisfies the repaired system: small residual everywhere
    res = o.residual_of(DT, p)
    assert np.abs(res[act]).max() < 1e-3 * np.abs(d1[act]).max()


def test_sealed_region_untouched_when_compatible():
    o, r = sealed_scene(OracleInstance)
    d = o.get("DIV")
    af = o.get("AREA_FRACTION")
    ls = o.get("LS_AIR0")
    act = (ls < 0) & ~(af == 0).all(-1)
    pocket = act & (r < 9.0)
    d[pocket] = 0.0  # compatible: the solve converges without the repair
    o.set("DIV", d)
    o.run_stage("PROJECTION", DT)
    assert o.stats().converged == 1
    assert np.array_equal(o.get("DIV"), d)
