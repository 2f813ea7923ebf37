"""The golden step vectors (tests/golden/): the CPU oracle reproduces them
bit for bit (a regression pin of the oracle), and the NumPy restatement of
the area fractions and the divergence agrees with them."""
import numpy as np
import pytest

from oracle.oracle import OracleInstance
from tests.golden_util import FILES, load, replay
from tests.test_oracle_numpy import np_area_fractions, np_divergence


def _exact(st, f, got, exp):
    assert got.dtype == exp.dtype and got.shape == exp.shape, (st, f)
    same = np.array_equal(got.view(np.uint8), exp.view(np.uint8))
    assert same, f"{st}: {f} differs (max |d| {np.abs(got.astype(np.float64) - exp).max()})"


@pytest.mark.parametrize("name", FILES)
def test_oracle_reproduces_golden_step(name):
    sc, state, stages = load(name)
    o = OracleInstance(sc["settings"])
    o.set_sources(sc["init_sources"], True)
    o.set_sources(sc["sources"], False)
    try:
        replay(o, o.set_obstacles, sc, state, stages, _exact)
    finally:
        o.close()


@pytest.mark.parametrize("name", FILES)
def test_numpy_restatement_matches_golden(name):
    sc, state, stages = load(name)
    exp = dict(stages)
    W, H = sc["settings"].width, sc["settings"].height
    cw, ch = 8 * (W // 8), 8 * (H // 8)  # the cell kernels' dispatch coverage (pipeline.rs:176-178)
    lss = exp["UPDATE_SOLID"]["LS_SOLID"]
    np.testing.assert_array_equal(np_area_fractions(lss)[:ch, :cw], exp["AREA_FRACTION"]["AREA_FRACTION"][:ch, :cw])
    u1 = exp["APPLY_FORCES"].get("U1", exp["ADVECT_VELOCITY"]["U1"])
    v1 = exp["APPLY_FORCES"].get("V1", exp["ADVECT_VELOCITY"]["V1"])
    us, vs = exp["UPDATE_SOLID"]["U_SOLID"], exp["UPDATE_SOLID"]["V_SOLID"]
    dx = 1.0 / sc["settings"].length_unit
    div = np_divergence(u1, v1, us, vs, lss, dx)
    np.testing.assert_array_equal(div[:ch, :cw], exp["DIVERGENCE"]["DIV"][:ch, :cw])
