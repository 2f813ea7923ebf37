"""ctypes mirror of include/efl.h: structs, enums and status codes.

The C ABI is the drop-in boundary of the per-step pipeline; these types are
what a Python host (and the tests) pass across it.  Layouts are checked
against the C sizes in tests/test_abi.py.
"""
import ctypes as C

ABI_VERSION = 1
MAX_FLUID_SOURCES = 16
MAX_SOLIDS = 256
MAX_BATCH = 64

OK = 0
E_INVALID = -1
E_HIP = -2
E_OOM = -3
E_STATE = -4
E_RANGE = -5
E_NOT_CONVERGED = -6

STATUS_NAMES = {
    OK: "EFL_OK", E_INVALID: "EFL_E_INVALID", E_HIP: "EFL_E_HIP", E_OOM: "EFL_E_OOM",
    E_STATE: "EFL_E_STATE", E_RANGE: "EFL_E_RANGE", E_NOT_CONVERGED: "EFL_E_NOT_CONVERGED",
}

# efl_projection_kind
PROJECTION_JACOBI = 0
PROJECTION_GAUSS_SEIDEL = 1
PROJECTION_MULTIGRID = 2
PROJECTION_MGPCG = 3
PROJECTION_LEARNED = 4

# efl_reinit_kind
REINIT_JUMP_FLOODING = 0
REINIT_FAST_ITERATIVE = 1

# efl_field
FIELD_NAMES = [
    "U0", "U1", "V0", "V1", "U_SOLID", "V_SOLID", "SOLID_ID", "U_VALID0", "U_VALID1",
    "V_VALID0", "V_VALID1", "DIV", "P0", "P1", "LS_AIR0", "LS_AIR1", "GRAD_LS", "LS_SOLID",
    "AREA_FRACTION", "FIM_LABELS0", "FIM_LABELS", "JFA_SEEDS0", "JFA_SEEDS1", "SOLID_BINS",
    "SOLID_FORCES", "MG_R0",
]
FIELD = {n: i for i, n in enumerate(FIELD_NAMES)}
FIELD_COUNT = len(FIELD_NAMES)
INT_FIELDS = {"SOLID_ID", "U_VALID0", "U_VALID1", "V_VALID0", "V_VALID1", "SOLID_BINS"}
UINT_FIELDS = {"FIM_LABELS0", "FIM_LABELS"}

# efl_stage
STAGE_NAMES = [
    "INIT_GRID_CENTER", "INIT_GRID_EDGE", "SOURCE_ONESHOT", "UPDATE_SOLID", "AREA_FRACTION",
    "ADVECT_VELOCITY", "APPLY_FORCES", "DIVERGENCE", "PROJECTION", "SOLVE_VELOCITY",
    "EXTRAPOLATE_VELOCITY", "ADVECT_LEVELSET", "SOURCE", "REINITIALIZE", "LEVELSET_GRADIENT",
    "FLUID_TO_SOLID", "GS_SWEEP", "MG_VCYCLE", "MG_RESIDUAL", "MG_SMOOTH", "GS_SWEEP_SPLIT",
    "NN_PRECONDITION", "MG_BOTTOM", "CG_SMOOTH",
]
STAGE = {n: i for i, n in enumerate(STAGE_NAMES)}
STAGE_COUNT = len(STAGE_NAMES)


class efl_projection(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("gs_iterations", C.c_uint32),
        ("mg_pre_iterations", C.c_uint32),
        ("mg_post_iterations", C.c_uint32),
        ("mg_coarsest_iterations", C.c_uint32),
        ("max_cycles", C.c_uint32),
        ("rel_tol", C.c_float),
        ("check_every", C.c_uint32),
    ]


class efl_reinit(C.Structure):
    _fields_ = [("kind", C.c_int32), ("fim_iterations", C.c_uint32)]


class efl_settings(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("rho", C.c_float),
        ("gravity", C.c_float * 2),
        ("length_unit", C.c_float),
        ("transform", C.c_float * 16),
        ("projection", efl_projection),
        ("reinit", efl_reinit),
    ]


class efl_source(C.Structure):
    _fields_ = [
        ("center", C.c_float * 2),
        ("data", C.c_float * 2),
        ("velocity", C.c_float * 2),
        ("shape", C.c_uint32),
        ("mode", C.c_uint32),
    ]


class efl_force(C.Structure):
    _fields_ = [("force", C.c_float * 2), ("position", C.c_float * 2)]


class efl_obstacle(C.Structure):
    _fields_ = [
        ("entity_id", C.c_uint32),
        ("shape", C.c_uint32),
        ("values", C.c_float * 6),
        ("transform", C.c_float * 16),
        ("inverse_transform", C.c_float * 16),
        ("linear_velocity", C.c_float * 2),
        ("angular_velocity", C.c_float),
        ("reserved", C.c_float),
    ]


class efl_body(C.Structure):
    _fields_ = [("density", C.c_float), ("area", C.c_float), ("inertia", C.c_float), ("dynamic", C.c_uint32)]


class efl_solid_force(C.Structure):
    _fields_ = [("force", C.c_float * 2), ("torque", C.c_float), ("torque_world", C.c_float)]


class efl_stats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("cycles", C.c_uint32),
        ("converged", C.c_int32),
        ("rel_residual", C.c_float),
        ("max_norm_divergence", C.c_float),
        ("projection_ms", C.c_float),
        ("step_ms", C.c_float),
        ("unconverged_solves", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class efl_fluid_diagnostics(C.Structure):
    _fields_ = [
        ("volume", C.c_float),
        ("min_velocity", C.c_float),
        ("max_velocity", C.c_float),
        ("reserved", C.c_float),
    ]


class efl_tuning(C.Structure):
    _fields_ = [
        ("pre", C.c_uint32),
        ("post", C.c_uint32),
        ("coarsest", C.c_uint32),
        ("bottom_cells", C.c_uint32),
        ("nn_strips", C.c_uint32),
        ("tile_h", C.c_uint32),
        ("trials", C.c_uint32),
        ("best_ms", C.c_float),
        ("default_ms", C.c_float),
    ]


STRUCT_SIZES = {
    "efl_projection": 32, "efl_reinit": 8, "efl_settings": 128, "efl_source": 32,
    "efl_force": 16, "efl_obstacle": 176, "efl_body": 16, "efl_solid_force": 16, "efl_stats": 40,
    "efl_tuning": 36, "efl_fluid_diagnostics": 16,
}


def struct_array(cls, items):
    """A ctypes array of `cls` built from a list of instances (or None if empty)."""
    n = len(items)
    if n == 0:
        return None, 0
    arr = (cls * n)(*items)
    return arr, n
