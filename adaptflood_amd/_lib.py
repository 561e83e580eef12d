"""ctypes binding of the native library (include/adaptflood.h).

The library is built in-tree by adaptflood_amd/build.py.  Loading fails loudly
when it is missing: there is no Python or CPU fallback for any entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libadaptflood_amd.so")

AF_OK = 0
AF_FP32 = 0
AF_FP64 = 1
AF_BC_CLOSED, AF_BC_FREE, AF_BC_HFIX, AF_BC_HVAR, AF_BC_QFIX, AF_BC_QVAR = range(6)
AF_FLOODPLAINQ_WORKSPACE_BYTES = 256
AF_SCN_OK, AF_SCN_NONFINITE, AF_SCN_STALLED = 0, 1, 2
AF_E_INVALID, AF_E_CAPACITY, AF_E_STATE = -1, -4, -5

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class ScenarioDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("dx", C.c_double), ("nodata", C.c_double),
        ("dem", _dp),
        ("fpn", C.c_double), ("depth_thresh", C.c_double), ("max_hflow", C.c_double),
        ("init_tstep", C.c_double), ("cfl", C.c_double), ("g", C.c_double),
        ("theta", C.c_double), ("sim_time", C.c_double), ("mass_int", C.c_double),
        ("west_free", C.c_int32), ("east_free", C.c_int32),
        ("north_free", C.c_int32), ("south_free", C.c_int32),
        ("n_points", C.c_int32),
        ("pt_x", _ip), ("pt_y", _ip), ("pt_type", _ip), ("pt_value", _dp),
        ("pt_series_off", _ip), ("pt_series_len", _ip),
        ("n_series", C.c_int32), ("series_t", _dp), ("series_v", _dp),
        ("rain_enabled", C.c_int32), ("n_rain", C.c_int32),
        ("rain_t", _dp), ("rain_r", _dp),
        ("h0", _dp),
    ]


class RunInfo(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("steps_done", C.c_int64), ("last_dt", C.c_double),
        ("n_reports", C.c_int32), ("finished", C.c_int32), ("hmax", C.c_double),
        ("status", C.c_int32), ("reserved", C.c_int32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("tile_rows", C.c_int32), ("waves_per_block", C.c_int32), ("graph_steps", C.c_int32),
        ("dt_log_cap", C.c_int32), ("report_cap", C.c_int32), ("variant", C.c_int32),
        ("tile_order", C.c_int32),
    ]


class Par(C.Structure):
    _fields_ = [
        ("dem_file", C.c_char * 512), ("bci_file", C.c_char * 512),
        ("bdy_file", C.c_char * 512), ("rain_file", C.c_char * 512),
        ("stage_file", C.c_char * 512), ("output_dir", C.c_char * 512),
        ("sim_time", C.c_double), ("save_int", C.c_double), ("mass_int", C.c_double),
        ("stage_int", C.c_double), ("init_tstep", C.c_double), ("cfl", C.c_double),
        ("fpn", C.c_double), ("depth_thresh", C.c_double), ("nodata", C.c_double),
        ("output_format", C.c_int32), ("voutput", C.c_int32), ("async_output", C.c_int32),
    ]


# (name, restype, argtypes) of every exported symbol declared in the header
_V = C.c_void_p
_D = C.c_double
_I = C.c_int
_I32 = C.c_int32
_I64 = C.c_int64
SIGNATURES = [
    ("af_last_error", C.c_char_p, []),
    ("af_version", C.c_char_p, []),
    ("af_UpdateQs_GPU", _I, [_V, _V, _V, _V, _I, _I, _D, _V]),
    ("af_FloodplainQ_GPU", _I, [_V, _V, _V, _V, _V, _V, _I, _I, _D, _D, _D, _D, _D, _D, _dp,
                                _D, _D, _D, _D, _D, _D, _D, _V, _V]),
    ("af_FloodplainQ_GPU_noalloc", _I, [_V, _V, _V, _V, _V, _V, _I, _I, _D, _D, _D, _D, _D, _D,
                                        _dp, _D, _D, _D, _D, _D, _D, _D, _V]),
    ("af_BCs_GPU", _I, [_V, _V, _V, _V, _I, _I, _D, _D, _D, _D, _I, _I, _I, _I, _V]),
    ("af_ApplyPointBCs_GPU", _I, [_V, _V, _V, _V, _I, _I, _D, _D, _D, _V]),
    ("af_UpdateH_GPU", _I, [_V, _V, _V, _I, _I, _D, _D, _V]),
    ("af_Rainfall_GPU", _I, [_V, _V, _I, _I, _D, _D, _D, _V]),
    ("af_AddFlowBCs_GPU", _I, [_V, _V, _V, _I, _V]),
    ("af_ComputeVelocity_GPU", _I, [_V, _V, _V, _V, _V, _V, _I, _I, _D, _D, _V]),
    ("af_domain_arena_bytes", _I, [C.POINTER(ScenarioDesc), _I, C.POINTER(Tuning),
                                   C.POINTER(C.c_size_t)]),
    ("af_domain_create", _I, [C.POINTER(ScenarioDesc), _I, C.POINTER(Tuning), _V, C.c_size_t,
                              _V, C.POINTER(_V)]),
    ("af_domain_destroy", _I, [_V]),
    ("af_domain_advance", _I, [_V, _I64, _I, _V]),
    ("af_domain_run", _I, [_V, _I64, _I, _V]),
    ("af_domain_info", _I, [_V, C.POINTER(RunInfo), _V]),
    ("af_domain_read", _I, [_V, _I, _dp, _V]),
    ("af_domain_read_dt_log", _I, [_V, _dp, _I64, C.POINTER(_I64), _V]),
    ("af_domain_read_reports", _I, [_V, _dp, _I32, C.POINTER(_I32), _V]),
    ("af_domain_volume", _I, [_V, _dp, _dp, _V]),
    ("af_domain_state_ptr", _I, [_V, _I, C.POINTER(_V), C.POINTER(_I64)]),
    ("af_batch_arena_bytes", _I, [C.POINTER(ScenarioDesc), _I, _I, C.POINTER(Tuning),
                                  C.POINTER(C.c_size_t)]),
    ("af_batch_create", _I, [C.POINTER(ScenarioDesc), _I, _I, C.POINTER(Tuning), _V,
                             C.c_size_t, _V, C.POINTER(_V)]),
    ("af_batch_size", _I, [_V, C.POINTER(_I), C.POINTER(_I)]),
    ("af_batch_finished", _I, [_V, C.POINTER(_I), _V]),
    ("af_batch_info", _I, [_V, _I, C.POINTER(RunInfo), _V]),
    ("af_batch_set_sim_time", _I, [_V, _I, C.c_double, _V]),
    ("af_batch_read_cells", _I, [_V, _I, C.POINTER(_I64), _I, _dp, _V]),
    ("af_batch_read", _I, [_V, _I, _I, _dp, _V]),
    ("af_batch_snapshot", _I, [_V, _I, _V, _V, _V, _V]),
    ("af_batch_read_dt_log", _I, [_V, _I, _dp, _I64, C.POINTER(_I64), _V]),
    ("af_batch_read_reports", _I, [_V, _I, _dp, _I32, C.POINTER(_I32), _V]),
    ("af_batch_volume", _I, [_V, _I, _dp, _dp, _V]),
    ("af_batch_launch_step", _I, [_V, _V]),
    ("af_pool_arena_bytes", _I, [_I, _I, C.c_size_t, C.POINTER(C.c_size_t)]),
    ("af_pool_scenario_bytes", _I, [C.POINTER(ScenarioDesc), _I, C.POINTER(Tuning),
                                    C.POINTER(C.c_size_t), _ip]),
    ("af_pool_create", _I, [_I, C.POINTER(Tuning), _I, _I, _V, C.c_size_t, C.POINTER(_V)]),
    ("af_pool_admit", _I, [_V, C.POINTER(ScenarioDesc), _V, _ip]),
    ("af_pool_stage", _I, [C.POINTER(ScenarioDesc), _I, C.POINTER(_V)]),
    ("af_pool_admit_staged", _I, [_V, _V, _V, _ip]),
    ("af_staged_free", None, [_V]),
    ("af_pool_release", _I, [_V, _I32, _V]),
    ("af_pool_poll", _I, [_V, _ip, _I32, _ip, _ip, _V]),
    ("af_pool_progress", _I, [_V, _ip, _dp, C.POINTER(_I64), _ip, _ip, _I32, _ip, _V]),
    ("af_pool_set_lane", _I, [_V, _ip, _I32, _V]),
    ("af_pool_advance_lane", _I, [_V, _I64, _I, _V]),
    ("af_pool_lane_ms", _I, [_V, C.POINTER(C.c_float)]),
    ("af_pool_stats", _I, [_V, _ip, _ip, C.POINTER(C.c_size_t), C.POINTER(_I64)]),
    ("af_domain_variant", _I, [_V, _ip]),
    ("af_par_defaults", None, [C.POINTER(Par)]),
    ("af_read_par", _I, [C.c_char_p, C.POINTER(Par)]),
    ("af_case_new", _V, []),
    ("af_case_free", None, [_V]),
    ("af_case_read_dem", _I, [_V, C.c_char_p]),
    ("af_case_read_bci", _I, [_V, C.c_char_p]),
    ("af_case_read_bdy", _I, [_V, C.c_char_p]),
    ("af_case_read_rain", _I, [_V, C.c_char_p]),
    ("af_case_load", _I, [_V, C.c_char_p]),
    ("af_case_load_par", _I, [_V, C.POINTER(Par)]),
    ("af_case_scenario", _I, [_V, C.POINTER(ScenarioDesc)]),
    ("af_case_par", _I, [_V, C.POINTER(Par)]),
    ("af_case_dem_header", _I, [_V, _ip, _ip, _dp, _dp, _dp, _dp]),
    ("af_case_line_bcs", _I, [_V, C.c_char_p, _I32, _ip]),
    ("af_case_point", _I, [_V, _I32, _ip, _ip, _ip, _dp, _ip, C.c_char_p, _I32]),
    ("af_case_point_series", _I, [_V, _I32, _dp, _dp, _I32]),
    ("af_case_rain", _I, [_V, _ip, _dp, _dp, _I32, _ip]),
    ("af_case_read_stage", _I, [_V, C.c_char_p]),
    ("af_case_stage", _I, [_V, _I32, _ip, _ip, C.c_char_p, _I32, _ip]),
    ("af_interp_series", _D, [_dp, _dp, _I32, _D, _D]),
    ("af_write_asc", _I, [C.c_char_p, _dp, _I32, _I32, _D, _D, _D]),
    ("af_write_asc_async", _I, [C.c_char_p, _dp, _I32, _I32, _D, _D, _D, _D]),
    ("af_nc_create", _I, [C.c_char_p, _I32, _I32, _D, _D, _D, _D, _I32, C.POINTER(_V)]),
    ("af_nc_append", _I, [_V, _D, _dp, _dp, _dp]),
    ("af_nc_close", _I, [_V]),
]

_lib = None


class AFError(RuntimeError):
    """A native entry point returned a non-zero status."""

    def __init__(self, name, code, msg):
        super().__init__(f"{name} returned {code}: {msg}")
        self.code = code


def lib():
    """Loads the native library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"native library missing: {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python adaptflood_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(name, rc):
    if rc != AF_OK:
        msg = lib().af_last_error().decode(errors="replace")
        raise AFError(name, rc, msg)
    return rc


def call(name, *args):
    """Calls entry point `name` and raises AFError on a non-zero status."""
    return check(name, getattr(lib(), name)(*args))


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)
