"""efl_fluid_diagnostics_noalloc against efl_get_fluid_diagnostics.

The allocation-free entry point runs the same kernels on caller-owned device
fields with the library's static per-GPU scratch for the block partials.  Two
calls in a row, the second on a smaller grid (so it reads only partials it
wrote itself), must give the instance entry point's values bit for bit; a grid
beyond the scratch is refused before anything is launched.  The device output
buffer is the caller's: the test allocates it with the HIP runtime.
"""
import ctypes as C

import numpy as np
import pytest

from bevy_eulerian_fluid_amd import abi, native
from tests.parity_util import small_scene

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


class _Hip:
    """The few HIP runtime calls the test needs for its own output buffer."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), n) == 0
        return p.value

    def to_host(self, dptr, n_floats):
        out = np.zeros(n_floats, np.float32)
        assert self.L.hipMemcpy(out.ctypes.data, C.c_void_p(dptr), 4 * n_floats, 2) == 0  # device to host
        return out

    def free(self, p):
        self.L.hipFree(C.c_void_p(p))


def _stepped(ctx, W, H, steps=4):
    sc = small_scene(W, H)
    i = ctx.instance(sc["settings"])
    i.set_sources(sc["init_sources"], True)
    i.set_sources(sc["sources"], False)
    ctx.set_obstacles(sc["obstacles"])
    for _ in range(steps):
        i.set_forces(sc["forces"])
        i.step(DT)
    return i


def test_noalloc_diagnostics_match_instance_entry_point():
    print("no-alloc fluid diagnostics: 512x384 then 200x136", flush=True)
    hip = _Hip()
    ctx = native.Context(0)
    out = hip.malloc(16)
    insts = []
    try:
        for (W, H) in [(512, 384), (200, 136)]:  # the second call is smaller
            i = _stepped(ctx, W, H)
            insts.append(i)
            want = i.fluid_diagnostics()
            ctx.fluid_diagnostics_noalloc(i.device_ptr("LS_AIR0"), i.device_ptr("U0"), i.device_ptr("V0"), W, H, out)
            ctx.sync()
            got = hip.to_host(out, 3)
            exp = np.array(want, np.float32)
            assert np.array_equal(got, exp), (W, H, got, exp)
            assert got[0] > 0 and got[2] > 0
        # beyond the scratch: refused with EFL_E_RANGE before any launch
        i = insts[0]
        with pytest.raises(native.EflError) as e:
            ctx.fluid_diagnostics_noalloc(i.device_ptr("LS_AIR0"), i.device_ptr("U0"), i.device_ptr("V0"),
                                          8192, 8192, out)
        assert e.value.code == abi.E_RANGE
    finally:
        for i in insts:
            i.close()
        hip.free(out)
        ctx.close()
