"""ctypes binding of libefl.so (the HIP library behind include/efl.h).

This module is the only path from Python to the fluid pipeline: there is no
CPU fallback.  If libefl.so is missing or cannot be loaded the import of a
context fails with NativeLibraryError naming the build command.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libefl.so")
_lib = None

EXPORTS = [
    "efl_ctx_create", "efl_ctx_destroy", "efl_last_error", "efl_sync", "efl_set_timing", "efl_set_graphs",
    "efl_set_obstacles", "efl_instance_create", "efl_instance_destroy", "efl_set_projection",
    "efl_set_reinit", "efl_set_uniforms", "efl_set_sources", "efl_set_forces", "efl_init_levelset",
    "efl_reset", "efl_initialize", "efl_step", "efl_run_stage", "efl_field_info", "efl_get_field",
    "efl_set_field", "efl_field_device_ptr", "efl_get_solid_forces", "efl_set_diagnostics",
    "efl_measure", "efl_get_stats", "efl_time_stage", "efl_set_bodies", "efl_get_obstacles",
    "efl_set_instance_obstacles", "efl_get_instance_obstacles",
    "efl_nn_param_count", "efl_nn_flops_per_pixel", "efl_nn_create", "efl_nn_destroy", "efl_nn_init_params",
    "efl_nn_set_params", "efl_nn_get_params", "efl_nn_dataset_reserve", "efl_nn_add_sample",
    "efl_nn_sample_count", "efl_nn_get_sample", "efl_nn_set_loss", "efl_nn_train", "efl_nn_eval", "efl_set_learned",
    "efl_nn_last_solve", "efl_autotune", "efl_apply_tuning", "efl_get_fluid_diagnostics", "efl_fluid_diagnostics_noalloc",
    "efl_set_async", "efl_host_waits",
]


class NativeLibraryError(RuntimeError):
    pass


class EflError(RuntimeError):
    def __init__(self, code, what, detail=""):
        self.code = code
        super().__init__(f"{what} failed: {abi.STATUS_NAMES.get(code, code)} {detail}".strip())


def load():
    """Load libefl.so and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            f"{LIB_PATH} is missing; build it with `make -C bevy_eulerian_fluid_amd/csrc` "
            "or __graft_entry__.build()")
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    P = C.c_void_p
    PP = C.POINTER(C.c_void_p)
    sig = {
        "efl_ctx_create": [C.c_int, PP],
        "efl_ctx_destroy": [P],
        "efl_sync": [P],
        "efl_set_timing": [P, C.c_int],
        "efl_set_graphs": [P, C.c_int],
        "efl_set_obstacles": [P, P, C.c_uint32],
        "efl_instance_create": [P, C.POINTER(abi.efl_settings), PP],
        "efl_instance_destroy": [P],
        "efl_set_projection": [P, C.POINTER(abi.efl_projection)],
        "efl_set_reinit": [P, C.POINTER(abi.efl_reinit)],
        "efl_set_uniforms": [P, C.c_float, P, P],
        "efl_set_sources": [P, P, C.c_uint32, C.c_int],
        "efl_set_forces": [P, P, C.c_uint32],
        "efl_init_levelset": [P, P],
        "efl_reset": [P],
        "efl_initialize": [P],
        "efl_step": [P, C.POINTER(C.c_void_p), C.c_uint32, C.c_float],
        "efl_run_stage": [P, C.c_int, C.c_float],
        "efl_field_info": [P, C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_uint32),
                           C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
        "efl_get_field": [P, C.c_int, P, C.c_size_t],
        "efl_set_field": [P, C.c_int, P, C.c_size_t],
        "efl_field_device_ptr": [P, C.c_int, PP],
        "efl_get_solid_forces": [P, P, C.c_uint32],
        "efl_set_diagnostics": [P, C.c_int],
        "efl_measure": [P, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "efl_get_stats": [P, C.POINTER(abi.efl_stats)],
        "efl_get_fluid_diagnostics": [P, C.POINTER(abi.efl_fluid_diagnostics)],
        "efl_time_stage": [P, C.c_int, C.c_float, C.c_uint32, C.POINTER(C.c_float)],
        "efl_set_bodies": [P, P, C.c_uint32],
        "efl_get_obstacles": [P, P, C.c_uint32],
        "efl_set_instance_obstacles": [P, P, P, C.c_uint32],
        "efl_get_instance_obstacles": [P, P, C.c_uint32],
        "efl_nn_create": [P, PP],
        "efl_nn_destroy": [P],
        "efl_nn_init_params": [P, C.c_uint64],
        "efl_nn_set_params": [P, P, C.c_size_t],
        "efl_nn_get_params": [P, P, C.c_size_t],
        "efl_nn_dataset_reserve": [P, C.c_uint32, C.c_uint32, C.c_uint32],
        "efl_nn_add_sample": [P, P, P, P],
        "efl_nn_sample_count": [P, C.POINTER(C.c_uint32)],
        "efl_nn_get_sample": [P, C.c_uint32, P],
        "efl_nn_set_loss": [P, C.c_int],
        "efl_nn_train": [P, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, P],
        "efl_nn_eval": [P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float), P, P],
        "efl_set_learned": [P, P],
        "efl_nn_last_solve": [P, P, C.POINTER(C.c_float)],
        "efl_autotune": [P, C.c_float, C.c_int, C.POINTER(abi.efl_tuning)],
        "efl_apply_tuning": [P, C.POINTER(abi.efl_tuning)],
        "efl_fluid_diagnostics_noalloc": [P, P, P, P, C.c_uint32, C.c_uint32, P],
        "efl_set_async": [P, C.c_uint32],
        "efl_host_waits": [P, C.POINTER(C.c_uint64)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.efl_last_error.argtypes = [P]
    L.efl_last_error.restype = C.c_char_p
    L.efl_nn_param_count.argtypes = []
    L.efl_nn_param_count.restype = C.c_size_t
    L.efl_nn_flops_per_pixel.argtypes = []
    L.efl_nn_flops_per_pixel.restype = C.c_double
    _lib = L
    return L


class Context:
    """One efl_ctx: a device, a stream and the global obstacle list."""

    def __init__(self, device=0):
        L = load()
        h = C.c_void_p()
        rc = L.efl_ctx_create(int(device), C.byref(h))
        if rc != abi.OK:
            raise EflError(rc, "efl_ctx_create", "(no HIP device?)")
        self.h = h
        self.L = L
        self._instances = []

    def _chk(self, rc, what):
        if rc != abi.OK:
            detail = self.L.efl_last_error(self.h)
            raise EflError(rc, what, detail.decode() if detail else "")

    def close(self):
        if self.h:
            self.L.efl_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        self._chk(self.L.efl_sync(self.h), "efl_sync")

    def set_timing(self, on):
        self._chk(self.L.efl_set_timing(self.h, int(bool(on))), "efl_set_timing")

    def set_graphs(self, on):
        """Replay steps as HIP graphs, or launch every kernel eagerly (default)."""
        self._chk(self.L.efl_set_graphs(self.h, int(bool(on))), "efl_set_graphs")

    def set_async(self, margin):
        """efl_set_async: 0 = MGPCG steps read their stopping state back;
        margin > 0 = they queue the last iteration count + margin and return."""
        self._chk(self.L.efl_set_async(self.h, int(margin)), "efl_set_async")

    def host_waits(self):
        n = C.c_uint64()
        self._chk(self.L.efl_host_waits(self.h, C.byref(n)), "efl_host_waits")
        return n.value

    def set_obstacles(self, obstacles):
        arr, n = abi.struct_array(abi.efl_obstacle, list(obstacles))
        self._chk(self.L.efl_set_obstacles(self.h, arr, n), "efl_set_obstacles")

    def set_bodies(self, bodies):
        arr, n = abi.struct_array(abi.efl_body, list(bodies))
        self._chk(self.L.efl_set_bodies(self.h, arr, n), "efl_set_bodies")

    def get_obstacles(self, n):
        out = (abi.efl_obstacle * n)()
        self._chk(self.L.efl_get_obstacles(self.h, out, n), "efl_get_obstacles")
        return list(out)

    def instance(self, settings):
        return Instance(self, settings)

    def fluid_diagnostics_noalloc(self, ls_air, u, v, width, height, out):
        """efl_fluid_diagnostics_noalloc on device pointers (ints): enqueues
        {volume, min, max} into the 3-float device buffer `out`; read it after
        sync().  Allocates nothing."""
        self._chk(self.L.efl_fluid_diagnostics_noalloc(self.h, C.c_void_p(ls_air), C.c_void_p(u), C.c_void_p(v),
                                                       int(width), int(height), C.c_void_p(out)),
                  "efl_fluid_diagnostics_noalloc")

    def step(self, instances, dt):
        n = len(instances)
        arr = (C.c_void_p * n)(*[i.h for i in instances])
        self._chk(self.L.efl_step(self.h, arr, n, C.c_float(dt)), "efl_step")


class Instance:
    """One efl_instance: the grid buffer set of one fluid entity."""

    def __init__(self, ctx, settings):
        self.ctx = ctx
        self.L = ctx.L
        h = C.c_void_p()
        ctx._chk(self.L.efl_instance_create(ctx.h, C.byref(settings), C.byref(h)), "efl_instance_create")
        self.h = h
        self.W = settings.width
        self.H = settings.height

    def close(self):
        if self.h and self.ctx.h:
            self.L.efl_instance_destroy(self.h)
        self.h = None

    def _chk(self, rc, what):
        self.ctx._chk(rc, what)

    def set_projection(self, p):
        self._chk(self.L.efl_set_projection(self.h, C.byref(p)), "efl_set_projection")

    def set_reinit(self, r):
        self._chk(self.L.efl_set_reinit(self.h, C.byref(r)), "efl_set_reinit")

    def set_uniforms(self, rho, gravity, transform=None):
        g = (C.c_float * 2)(*gravity)
        t = (C.c_float * 16)(*transform) if transform is not None else None
        self._chk(self.L.efl_set_uniforms(self.h, C.c_float(rho), g, t), "efl_set_uniforms")

    def set_sources(self, sources, oneshot):
        arr, n = abi.struct_array(abi.efl_source, list(sources))
        self._chk(self.L.efl_set_sources(self.h, arr, n, int(bool(oneshot))), "efl_set_sources")

    def set_forces(self, forces):
        arr, n = abi.struct_array(abi.efl_force, list(forces))
        self._chk(self.L.efl_set_forces(self.h, arr, n), "efl_set_forces")

    def set_obstacles(self, obstacles, bodies=None):
        """This instance's own obstacle list (None = use the context's)."""
        if obstacles is None:
            self._chk(self.L.efl_set_instance_obstacles(self.h, None, None, 0), "efl_set_instance_obstacles")
            return
        obs = list(obstacles)
        arr, n = abi.struct_array(abi.efl_obstacle, obs)
        if arr is None:
            arr = (abi.efl_obstacle * 1)()
        barr = None
        if bodies is not None:
            barr, _ = abi.struct_array(abi.efl_body, list(bodies))
            if barr is None:
                barr = (abi.efl_body * 1)()
        self._chk(self.L.efl_set_instance_obstacles(self.h, arr, barr, n), "efl_set_instance_obstacles")

    def get_obstacles(self, n):
        out = (abi.efl_obstacle * n)()
        self._chk(self.L.efl_get_instance_obstacles(self.h, out, n), "efl_get_instance_obstacles")
        return list(out)

    def init_levelset(self, ls):
        if ls is None:
            self._chk(self.L.efl_init_levelset(self.h, None), "efl_init_levelset")
            return
        a = np.ascontiguousarray(ls, dtype=np.float32).reshape(-1)
        assert a.size == self.W * self.H
        self._chk(self.L.efl_init_levelset(self.h, a.ctypes.data), "efl_init_levelset")

    def reset(self):
        self._chk(self.L.efl_reset(self.h), "efl_reset")

    def initialize(self):
        self._chk(self.L.efl_initialize(self.h), "efl_initialize")

    def step(self, dt):
        self.ctx.step([self], dt)

    def run_stage(self, stage, dt):
        if isinstance(stage, str):
            stage = abi.STAGE[stage]
        self._chk(self.L.efl_run_stage(self.h, int(stage), C.c_float(dt)), "efl_run_stage")

    def set_diagnostics(self, on):
        self._chk(self.L.efl_set_diagnostics(self.h, int(bool(on))), "efl_set_diagnostics")

    def time_stage(self, stage, dt, reps):
        """Mean device ms per repetition of one stage (hipEvents on the
        library's stream)."""
        if isinstance(stage, str):
            stage = abi.STAGE[stage]
        ms = C.c_float()
        self._chk(self.L.efl_time_stage(self.h, int(stage), C.c_float(dt), int(reps), C.byref(ms)),
                  "efl_time_stage")
        return ms.value

    def field_info(self, field):
        if isinstance(field, str):
            field = abi.FIELD[field]
        n = C.c_size_t()
        w, h, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._chk(self.L.efl_field_info(self.h, int(field), C.byref(n), C.byref(w), C.byref(h), C.byref(c)),
                  "efl_field_info")
        return n.value, w.value, h.value, c.value

    def _dtype(self, name):
        if name in abi.INT_FIELDS:
            return np.int32
        if name in abi.UINT_FIELDS:
            return np.uint32
        return np.float32

    def get(self, field):
        name = field if isinstance(field, str) else abi.FIELD_NAMES[field]
        n, w, h, c = self.field_info(name)
        out = np.empty(n, dtype=self._dtype(name))
        self._chk(self.L.efl_get_field(self.h, abi.FIELD[name], out.ctypes.data, n), "efl_get_field")
        return out.reshape((h, w, c)) if c > 1 else out.reshape((h, w))

    def set(self, field, arr):
        name = field if isinstance(field, str) else abi.FIELD_NAMES[field]
        n, w, h, c = self.field_info(name)
        a = np.ascontiguousarray(arr, dtype=self._dtype(name)).reshape(-1)
        assert a.size == n, (name, a.size, n)
        self._chk(self.L.efl_set_field(self.h, abi.FIELD[name], a.ctypes.data, n), "efl_set_field")

    def device_ptr(self, field):
        if isinstance(field, str):
            field = abi.FIELD[field]
        p = C.c_void_p()
        self._chk(self.L.efl_field_device_ptr(self.h, int(field), C.byref(p)), "efl_field_device_ptr")
        return p.value

    def solid_forces(self, n=abi.MAX_SOLIDS):
        out = (abi.efl_solid_force * n)()
        self._chk(self.L.efl_get_solid_forces(self.h, out, n), "efl_get_solid_forces")
        return np.array([[f.force[0], f.force[1], f.torque] for f in out], dtype=np.float32)

    def measure(self, dt):
        r, d = C.c_float(), C.c_float()
        self._chk(self.L.efl_measure(self.h, C.c_float(dt), C.byref(r), C.byref(d)), "efl_measure")
        return r.value, d.value

    def autotune(self, dt, apply=True):
        """efl_autotune: the chosen settings and the solve times as a dict."""
        t = abi.efl_tuning()
        self._chk(self.L.efl_autotune(self.h, C.c_float(dt), int(bool(apply)), C.byref(t)), "efl_autotune")
        return {f: getattr(t, f) for f, _ in abi.efl_tuning._fields_}

    def apply_tuning(self, **kw):
        """efl_apply_tuning with the fields of an autotune() result; fields not
        given take the defaults (2, 2, 3, 1024, tile_h 0 = by level size, no
        64-pixel strips)."""
        t = abi.efl_tuning(pre=2, post=2, coarsest=3, bottom_cells=1024, tile_h=0)
        for k, val in kw.items():
            setattr(t, k, val)
        self._chk(self.L.efl_apply_tuning(self.h, C.byref(t)), "efl_apply_tuning")

    def set_learned(self, net):
        self._chk(self.L.efl_set_learned(self.h, net.h if net is not None else None), "efl_set_learned")

    def learned_iterations(self, with_rel=False):
        a = (C.c_uint32 * 2)()
        rel = C.c_float()
        self._chk(self.L.efl_nn_last_solve(self.h, a, C.byref(rel)), "efl_nn_last_solve")
        return (int(a[0]), int(a[1]), rel.value) if with_rel else (int(a[0]), int(a[1]))

    def stats(self):
        s = abi.efl_stats()
        self._chk(self.L.efl_get_stats(self.h, C.byref(s)), "efl_get_stats")
        return s

    def fluid_diagnostics(self):
        """(volume, min speed, max speed) of the current fields, as the
        reference's diagnostics passes compute them."""
        d = abi.efl_fluid_diagnostics()
        self._chk(self.L.efl_get_fluid_diagnostics(self.h, C.byref(d)), "efl_get_fluid_diagnostics")
        return d.volume, d.min_velocity, d.max_velocity


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Network:
    """efl_nn: the learned preconditioner's parameters, training set and
    trainer on one context's device."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.L = load()
        h = C.c_void_p()
        ctx._chk(self.L.efl_nn_create(ctx.h, C.byref(h)), "efl_nn_create")
        self.h = h
        self.w = self.h_ = 0

    @staticmethod
    def param_count():
        return int(load().efl_nn_param_count())

    @staticmethod
    def flops_per_pixel():
        return float(load().efl_nn_flops_per_pixel())

    def close(self):
        if self.h:
            self.L.efl_nn_destroy(self.h)
            self.h = None

    def _chk(self, rc, what):
        self.ctx._chk(rc, what)

    def init_params(self, seed):
        self._chk(self.L.efl_nn_init_params(self.h, seed), "efl_nn_init_params")

    def set_params(self, p):
        a = _f32(p)
        self._chk(self.L.efl_nn_set_params(self.h, a.ctypes.data, a.size), "efl_nn_set_params")

    def get_params(self):
        a = np.zeros(self.param_count(), dtype=np.float32)
        self._chk(self.L.efl_nn_get_params(self.h, a.ctypes.data, a.size), "efl_nn_get_params")
        return a

    def reserve(self, w, h, capacity):
        self._chk(self.L.efl_nn_dataset_reserve(self.h, w, h, capacity), "efl_nn_dataset_reserve")
        self.w, self.h_ = w, h

    def add_sample(self, inst, rhs, target):
        r, t = _f32(rhs), _f32(target)
        self._chk(self.L.efl_nn_add_sample(self.h, inst.h, r.ctypes.data, t.ctypes.data), "efl_nn_add_sample")

    def count(self):
        n = C.c_uint32()
        self._chk(self.L.efl_nn_sample_count(self.h, C.byref(n)), "efl_nn_sample_count")
        return n.value

    def sample(self, index):
        a = np.zeros((5, self.h_, self.w), dtype=np.float32)
        self._chk(self.L.efl_nn_get_sample(self.h, index, a.ctypes.data), "efl_nn_get_sample")
        return a

    LOSS = {"mse": 0, "residual": 1}

    def set_loss(self, loss):
        """"mse" (output against the sample's target) or "residual" (the
        relative residual the best multiple of the output leaves)."""
        self._chk(self.L.efl_nn_set_loss(self.h, self.LOSS[loss]), "efl_nn_set_loss")

    def train(self, iterations, batch, lr, seed):
        hist = np.zeros(iterations, dtype=np.float32)
        self._chk(self.L.efl_nn_train(self.h, iterations, batch, lr, seed, hist.ctypes.data), "efl_nn_train")
        return hist

    def eval(self, first, count, grads=False, outputs=False):
        loss = C.c_float()
        g = np.zeros(self.param_count(), dtype=np.float32) if grads else None
        o = np.zeros((count, self.h_, self.w), dtype=np.float32) if outputs else None
        self._chk(self.L.efl_nn_eval(self.h, first, count, C.byref(loss),
                                     g.ctypes.data if grads else None, o.ctypes.data if outputs else None),
                  "efl_nn_eval")
        return loss.value, g, o
