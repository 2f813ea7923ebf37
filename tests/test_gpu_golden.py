"""The golden step vectors (tests/golden/) replayed through libefl.so, each
stage on its golden inputs: every field a stage changes must equal the
vector bit for bit, except the local forces' exp() (ocml vs glibc: 2e-6
relative, 1e-6 absolute on U1 / V1 after APPLY_FORCES) and the multigrid
residual scratch field, which the fused V-cycle does not store."""
import numpy as np
import pytest

from tests.golden_util import FILES, load, replay

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", FILES)
def test_hip_matches_golden_step(gpu_ctx, name):
    print(f"golden {name}", flush=True)
    sc, state, stages = load(name)
    inst = gpu_ctx.instance(sc["settings"])
    inst.set_sources(sc["init_sources"], True)
    inst.set_sources(sc["sources"], False)

    def compare(st, f, got, exp):
        assert got.shape == exp.shape and got.dtype == exp.dtype, (st, f)
        if f == "MG_R0":  # the fused V-cycle never stores the level-0 residual
            return
        if st == "APPLY_FORCES":
            np.testing.assert_allclose(got, exp, rtol=2e-6, atol=1e-6, err_msg=f"{st}: {f}")
            return
        assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), f"{st}: {f} differs"

    try:
        replay(inst, lambda obs: gpu_ctx.set_obstacles(obs), sc, state, stages, compare)
    finally:
        gpu_ctx.set_obstacles([])
        inst.close()
