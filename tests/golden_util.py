"""Loading the golden step vectors of tests/golden/ (written by
tests/golden/make_fixtures.py) and replaying them on an instance."""
import os

import numpy as np

from bevy_eulerian_fluid_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = ["step_64x64.npz", "step_72x68.npz"]
STAGES = ["UPDATE_SOLID", "AREA_FRACTION", "ADVECT_VELOCITY", "APPLY_FORCES", "DIVERGENCE", "PROJECTION",
          "SOLVE_VELOCITY", "EXTRAPOLATE_VELOCITY", "ADVECT_LEVELSET", "SOURCE", "REINITIALIZE",
          "LEVELSET_GRADIENT", "FLUID_TO_SOLID"]


def load(name):
    """(scene dict, {field: state}, [(stage, {field: expected})])"""
    import ctypes
    d = np.load(os.path.join(GOLDEN, name), allow_pickle=False)

    def structs(key, cls):
        raw = bytes(d[key])
        sz = ctypes.sizeof(cls)
        return [cls.from_buffer_copy(raw[k * sz:(k + 1) * sz]) for k in range(len(raw) // sz)]

    sc = {
        "settings": abi.efl_settings.from_buffer_copy(bytes(d["scene/settings"])),
        "init_sources": structs("scene/init_sources", abi.efl_source),
        "sources": structs("scene/sources", abi.efl_source),
        "obstacles": structs("scene/obstacles", abi.efl_obstacle),
        "obstacles_step": structs("scene/obstacles_step", abi.efl_obstacle),
        "forces": structs("scene/forces", abi.efl_force),
        "dt": float(d["scene/dt"][0]),
    }
    state = {k.split("/", 1)[1]: d[k] for k in d.files if k.startswith("state/")}
    stages = []
    for st in STAGES:
        exp = {k.split("/", 1)[1]: d[k] for k in d.files if k.startswith(st + "/")}
        stages.append((st, exp))
    return sc, state, stages


def replay(inst, set_obstacles, sc, state, stages, compare):
    """Load the state into inst, then run each stage and call
    compare(stage, field, got, expected) for every field it must change.
    Before each stage the fields earlier stages changed are reset to their
    golden values, so every stage is checked on exactly its golden inputs."""
    for f, a in state.items():
        inst.set(f, a)
    set_obstacles(sc["obstacles_step"])
    inst.set_forces(sc["forces"])
    golden = {}
    for st, exp in stages:
        for f, a in golden.items():
            inst.set(f, a)
        inst.run_stage(st, sc["dt"])
        for f, a in exp.items():
            compare(st, f, inst.get(f), a)
        golden.update(exp)
