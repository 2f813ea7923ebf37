"""A pressure system with a sealed region: fluid fills the grid below an air
layer, and an annulus of solid encloses a disc of fluid whose divergence
does not sum to zero (as when bodies close on the fluid between them)."""
import numpy as np

from bevy_eulerian_fluid_amd import scenes

DT = 1.0 / 60.0


def sealed_scene(Instance, W=64, H=64, seed=3):
    proj = scenes.projection("mgpcg", max_cycles=200, rel_tol=1e-5)
    s = scenes.settings(W, H, proj=proj, rein=scenes.reinit("fast_iterative", 10))
    o = Instance(s)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    r = np.hypot(xs - 0.5 * W + 0.3, ys - 0.55 * H)
    solid = np.abs(r - 11.0) - 3.2  # negative inside the ring
    o.set("LS_SOLID", solid.astype(np.float32))
    o.run_stage("AREA_FRACTION", DT)
    o.set("LS_AIR0", (6.5 - ys).astype(np.float32))  # air in the top rows
    rng = np.random.default_rng(seed)
    o.set("DIV", rng.uniform(-2.0, 3.0, (H, W)).astype(np.float32))
    o.set("P0", np.zeros((H, W), np.float32))
    return o, r
