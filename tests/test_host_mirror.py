"""The C++ host mirror of the reference's main-world API
(bevy_eulerian_fluid_amd/host/euler_fluid.hpp).

CPU: host_check's own checks of the pure host logic (FluidPlugin's panic,
Display strings and defaults, source/obstacle/readback conversions, body
moments against numerical integration).

GPU: a 128 x 96 fluid driven frame by frame through FluidWorld (oneshot and
continuous sources plus an inactive sink, five bodies one of which has an
unsupported shape, local forces on a stepping frame and on an idle frame).
The same frames are replayed through the C ABI directly from inputs restated
here from the reference's conversion rules (fluid_source_uniform.rs:77-158 -
inactive sources skipped, oneshot velocity zero; apply_forces.rs:157-167 - the
list of every frame replaces the last; render_node.rs:165-175 - no step
without a new physics step), and every field must agree bit for bit.  (The
kernels themselves are checked against the oracle elsewhere; over several
steps with local forces the two differ within the tolerance of exp.)
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

from bevy_eulerian_fluid_amd import abi, native, scenes
from tests.parity_util import ALL_FIELDS, compare

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_CHECK = os.path.join(ROOT, "bevy_eulerian_fluid_amd", "host", "host_check")


def _binary():
    if not os.path.exists(HOST_CHECK):
        pytest.fail("host_check is not built (run __graft_entry__.build())")
    return HOST_CHECK


def test_host_mirror_cpu_checks():
    out = subprocess.run([_binary(), "cpu"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = [ln for ln in out.stdout.splitlines() if ln.startswith(("ok", "FAIL"))]
    assert len(lines) >= 20 and all(ln.startswith("ok") for ln in lines), out.stdout
    # the unsupported shape is reported the way the reference warns
    assert "Unsupported shape type for solid" in out.stderr


def _read_records(path):
    recs = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        name = data[pos:pos + 32].split(b"\0", 1)[0].decode()
        (n,) = struct.unpack_from("<Q", data, pos + 32)
        recs[name] = data[pos + 40:pos + 40 + n]
        pos += 40 + n
    return recs


def _expected_obstacles():
    """The bodies of the scene as construct_rigid_body_buffer_for_gpu
    (obstacle.rs:89-160) turns them into records: ids are body indices (body
    2 has an unsupported shape and is skipped), Collider::rectangle takes full
    sizes."""
    return [
        scenes.obstacle(0, 0, [6.0], 0.0, (20.0, -10.0), lin_vel=(3.0, 1.0)),
        scenes.obstacle(1, 1, [8.0, 4.0], 0.5, (-25.0, -20.0), ang_vel=0.8),
        scenes.obstacle(3, 2, [-5.0, 0.0, 5.0, 0.0, 3.0], -0.3, (35.0, -25.0)),
        scenes.obstacle(4, 4, [-6.0, -4.0, 6.0, -4.0, 0.0, 8.0], 1.1, (0.0, 10.0)),
    ]


@pytest.mark.gpu
def test_host_mirror_scene_matches_direct_abi(tmp_path):
    print("host mirror: FluidWorld 128x96 scene through host_check", flush=True)
    dump = str(tmp_path / "scene.bin")
    out = subprocess.run([_binary(), "scene", dump], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    rec = _read_records(dump)

    frames = np.frombuffer(rec["frames"], dtype=np.int32).tolist()
    assert frames == [1, 0, 1, 0, 1, 1], frames
    assert np.frombuffer(rec["textures_ok"], dtype=np.int32)[0] == 1, "FluidTextures.u differs from field U0"

    n_obs = len(rec["obstacles"]) // C_sizeof_obstacle()
    obstacles = list((abi.efl_obstacle * n_obs).from_buffer_copy(rec["obstacles"]))
    expect = _expected_obstacles()
    assert n_obs == len(expect)
    for o, e in zip(obstacles, expect):
        assert (o.entity_id, o.shape) == (e.entity_id, e.shape)
        assert list(o.values) == list(e.values)
        np.testing.assert_allclose(list(o.transform), list(e.transform), rtol=0, atol=4e-6)
        np.testing.assert_allclose(list(o.inverse_transform), list(e.inverse_transform), rtol=0, atol=4e-5)
        assert list(o.linear_velocity) == list(e.linear_velocity) and o.angular_velocity == e.angular_velocity

    settings = scenes.settings(128, 96, rho=99.7, gravity=(0.0, 9.8), length_unit=10.0,
                               proj=scenes.projection("gauss_seidel", gs_iterations=20),
                               rein=scenes.reinit("fast_iterative", 10))
    ctx = native.Context(0)
    try:
        g = ctx.instance(settings)
        # update_fluid_source_init_buffer: the oneshot box with zero velocity
        g.set_sources([scenes.source((0.0, -16.0), "aabb", half_size=(64.0, 32.0))], True)
        # update_fluid_source_buffer: the continuous source; the sink is inactive
        g.set_sources([scenes.source((-48.0, 28.0), "aabb", half_size=(5.0, 5.0), velocity=(50.0, 10.0))], False)
        ctx.set_obstacles(obstacles)
        dt = np.float32(1.0 / 60.0).item()
        g.step(dt)  # frame 1: initialisation
        g.set_forces([scenes.force((0.0, -40.0), (64.0, 60.0)), scenes.force((25.0, 0.0), (30.0, 70.0))])
        g.step(dt)  # frame 3
        g.set_forces([])  # frame 4's force was replaced by frame 5's empty list
        g.step(dt)  # frame 5
        g.step(dt)  # frame 6

        errs = []
        for k, name in enumerate(ALL_FIELDS):
            ref = g.get(name)
            got = np.frombuffer(rec[f"field{k}"], dtype=ref.dtype).reshape(ref.shape)
            m = compare(got, ref, name)
            if m:
                errs.append(m)
        assert not errs, "\n".join(errs)
        forces = np.frombuffer(rec["forces_to_solid"], dtype=np.float32).reshape(-1, 3)
        ref = g.get("SOLID_FORCES").reshape(-1, 4)[: forces.shape[0], :3]
        assert forces.shape[0] == 5 and np.array_equal(forces, ref)
        assert np.abs(ref[[0, 1, 3, 4]]).max() > 0 and not ref[2].any()
        # the final frame ran the diagnostics passes (FluidDiagnosticsPlugin)
        # on the final state and left the fluid node idle
        dg = np.frombuffer(rec["diagnostics"], dtype=np.float32)
        assert dg[3] == 1.0, "the diagnostics frame advanced the fluid"
        assert np.array_equal(dg[:3], np.float32(g.fluid_diagnostics())) and dg[0] > 0
        g.close()
    finally:
        ctx.close()


def C_sizeof_obstacle():
    import ctypes
    return ctypes.sizeof(abi.efl_obstacle)


def test_expected_obstacles_are_rigid():
    """The restated records are proper rigid transforms (guards the
    restatement itself)."""
    for e in _expected_obstacles():
        m = np.array(list(e.transform), dtype=np.float64).reshape(4, 4).T
        inv = np.array(list(e.inverse_transform), dtype=np.float64).reshape(4, 4).T
        assert np.allclose(m @ inv, np.eye(4), atol=1e-5)
        assert math.isclose(m[0, 0] ** 2 + m[1, 0] ** 2, 1.0, rel_tol=1e-6)
