"""ctypes mirror of include/mgo.h and the loader of the device library.

The structs below follow include/mgo.h field by field. ``load()`` opens the in-tree
``memilio_amd/libmgo.so`` and raises if it is missing: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmgo.so")

NUM_COMPARTMENTS = 10
MAX_GROUPS = 8

# status codes (mirror mio::StatusCode where one exists)
OK = 0
ERR_INVALID_VALUE = 1
ERR_OUT_OF_RANGE = 2
ERR_OUT_OF_MEMORY = 3
ERR_UNSUPPORTED = 4
ERR_DEVICE = 5
ERR_NOT_FOUND = 6
ERR_BUSY = 7

# per-node scalar parameters
P_START_DAY = 0
P_SEASONALITY = 1
P_ICU_CAPACITY = 2
P_TEST_AND_TRACE_CAPACITY = 3
P_TEST_AND_TRACE_MAX_RISK = 4
P_COMMUTER_NONDETECTION = 5
P_START_COMMUTER_DETECTION = 6
P_END_COMMUTER_DETECTION = 7
NUM_SCALARS = 8

# per-node per-age parameters
A_TIME_EXPOSED = 0
A_TIME_INFECTED_NO_SYMPTOMS = 1
A_TIME_INFECTED_SYMPTOMS = 2
A_TIME_INFECTED_SEVERE = 3
A_TIME_INFECTED_CRITICAL = 4
A_TRANSMISSION_PROBABILITY_ON_CONTACT = 5
A_RELATIVE_TRANSMISSION_NO_SYMPTOMS = 6
A_RECOVERED_PER_INFECTED_NO_SYMPTOMS = 7
A_RISK_OF_INFECTION_FROM_SYMPTOMATIC = 8
A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC = 9
A_SEVERE_PER_INFECTED_SYMPTOMS = 10
A_CRITICAL_PER_SEVERE = 11
A_DEATHS_PER_SEVERE = 12
A_DEATHS_PER_CRITICAL = 13
NUM_AGE_PARAMS = 14

# integrator kinds
INTEGRATOR_CASH_KARP54 = 0
INTEGRATOR_RKF45 = 1
INTEGRATOR_CASH_KARP54_FIXED = 2
INTEGRATOR_EULER = 3

CTX_NO_GRAPH_CAPTURE = 1
CTX_PROFILE_KERNELS = 2
CTX_SEQUENTIAL_FIXUP = 4
SERVE_BATCH_KERNELS = 1  # mgo_serve_cfg.flags: per-(integrator, G) launches, bitwise equal to batch runs

DBL_MAX = np.finfo(np.float64).max
DBL_MIN = np.finfo(np.float64).tiny

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


NPI_DURATION, NPI_DELAY, NPI_BASE_VALUE, NPI_DIRECTIVE_BEGIN, NPI_DIRECTIVE_END = range(5)
NPI_NUM_PARAMS = 5
NPI_LIST_EXTRA = 32


class DynamicNpis(C.Structure):
    _fields_ = [
        ("node_offsets", _ip),
        ("node_params", _dp),
        ("thresholds", _dp),
        ("damping_offsets", _ip),
        ("damping_values", _dp),
        ("damping_levels", _ip),
        ("damping_types", _ip),
        ("damping_weights", _dp),
    ]


class GraphDesc(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32),
        ("num_groups", C.c_int32),
        ("node_scalars", _dp),
        ("age_params", _dp),
        ("contact_baseline", _dp),
        ("contact_minimum", _dp),
        ("damping_offsets", _ip),
        ("damping_times", _dp),
        ("damping_levels", _ip),
        ("damping_types", _ip),
        ("damping_coeffs", _dp),
        ("populations", _dp),
        ("num_edges", C.c_int32),
        ("edge_start", _ip),
        ("edge_end", _ip),
        ("edge_coeffs", _dp),
        ("edge_coeff_len", C.c_int32),
        ("dynamic_npis", C.POINTER(DynamicNpis)),
        ("edge_coeff_minimum", _dp),
        ("out_damping_offsets", _ip),
        ("out_damping_times", _dp),
        ("out_damping_levels", _ip),
        ("out_damping_types", _ip),
        ("out_damping_coeffs", _dp),
        ("out_dynamic_npis", C.POINTER(DynamicNpis)),
    ]


class IntegratorCfg(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("abs_tol", C.c_double),
        ("rel_tol", C.c_double),
        ("dt_min", C.c_double),
        ("dt_max", C.c_double),
    ]


class SimCfg(C.Structure):
    _fields_ = [
        ("t0", C.c_double),
        ("dt_graph", C.c_double),
        ("dt_node", C.c_double),
        ("horizon", C.c_double),
        ("integrator", IntegratorCfg),
    ]


class DeviceCfg(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("hbm_budget_bytes", C.c_size_t),
        ("flags", C.c_int32),
        ("result_pool_bytes", C.c_size_t),
    ]


class EnsembleDesc(C.Structure):
    _fields_ = [
        ("num_runs", C.c_int32),
        ("graphs", C.POINTER(GraphDesc)),
        ("cfgs", C.POINTER(SimCfg)),
        ("tmax", _dp),
    ]


class RunInfo(C.Structure):
    _fields_ = [
        ("device_ms", C.c_double),
        ("node_days", C.c_double),
        ("steps", C.c_int64),
        ("attempts", C.c_int64),
        ("exact_replays", C.c_int64),
        ("fixup_passes", C.c_int64),
        ("half_day_cycles", C.c_int32),
        ("node_launches", C.c_int32),
        ("node_kernel_ms", C.c_double),
        ("edge_launches", C.c_int32),
        ("edge_kernel_ms", C.c_double),
        ("tile", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Request(C.Structure):
    _fields_ = [
        ("graph", GraphDesc),
        ("cfg", SimCfg),
        ("tmax", C.c_double),
    ]


class ServeCfg(C.Structure):
    _fields_ = [
        ("max_resident", C.c_int32),
        ("hbm_budget_bytes", C.c_size_t),
        ("flags", C.c_int32),
        ("result_pool_bytes", C.c_size_t),
    ]


class DampingUpdate(C.Structure):
    _fields_ = [
        ("scale", C.c_double),
        ("t_start", C.c_double),
        ("level", C.c_int32),
        ("type", C.c_int32),
    ]


class ResultInfo(C.Structure):
    _fields_ = [
        ("ticket", C.c_int64),
        ("status", C.c_int32),
        ("submit_ms", C.c_double),
        ("admit_ms", C.c_double),
        ("retire_ms", C.c_double),
        ("done_ms", C.c_double),
        ("num_nodes", C.c_int32),
        ("num_groups", C.c_int32),
        ("num_days", C.c_int32),
    ]


class ServeStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("cycles", C.c_int64),
        ("admitted", C.c_int64),
        ("retired", C.c_int64),
        ("node_cycles", C.c_double),
        ("device_ms", C.c_double),
        ("host_ms", C.c_double),
        ("max_resident_seen", C.c_int32),
        ("reserved", C.c_int32),
    ]


# every symbol include/mgo.h declares, with its ctypes signature
SIGNATURES = {
    "mgo_context_create": (C.c_int32, [C.POINTER(DeviceCfg), C.POINTER(C.c_void_p)]),
    "mgo_context_destroy": (C.c_int32, [C.c_void_p]),
    "mgo_context_set_flags": (C.c_int32, [C.c_void_p, C.c_int32]),
    "mgo_context_set_tile": (C.c_int32, [C.c_void_p, C.c_int32]),
    "mgo_last_error": (C.c_char_p, [C.c_void_p]),
    "mgo_version": (C.c_char_p, []),
    "mgo_graph_create": (C.c_int32, [C.c_void_p, C.POINTER(GraphDesc), C.POINTER(SimCfg), C.POINTER(C.c_int64)]),
    "mgo_graph_advance": (C.c_int32, [C.c_void_p, C.c_int64, C.c_double]),
    "mgo_graph_get_time": (C.c_int32, [C.c_void_p, C.c_int64, _dp]),
    "mgo_graph_read_state": (C.c_int32, [C.c_void_p, C.c_int64, _dp]),
    "mgo_graph_read_daily": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, _dp, _ip]),
    "mgo_graph_read_stats": (C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mgo_graph_destroy": (C.c_int32, [C.c_void_p, C.c_int64]),
    "mgo_simulate_single": (C.c_int32, [C.c_void_p, C.POINTER(GraphDesc), C.POINTER(SimCfg), C.c_double, C.c_int32,
                                        _dp, _ip]),
    "mgo_ensemble_run": (C.c_int32, [C.c_void_p, C.POINTER(EnsembleDesc), _dp, C.POINTER(RunInfo)]),
    "mgo_batch_create": (C.c_int32, [C.c_void_p, C.POINTER(EnsembleDesc), C.POINTER(C.c_int64)]),
    "mgo_batch_run": (C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(RunInfo)]),
    "mgo_batch_read_daily": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, _dp]),
    "mgo_batch_reset": (C.c_int32, [C.c_void_p, C.c_int64]),
    "mgo_batch_destroy": (C.c_int32, [C.c_void_p, C.c_int64]),
    "mgo_batch_ensemble_mean": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, _dp]),
    "mgo_batch_ensemble_percentiles": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, _dp, C.c_int32, _dp]),
    "mgo_batch_sum_nodes": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, _dp]),
    "mgo_ensemble_mean_values": (C.c_int32, [C.c_void_p, _dp, C.c_int32, C.c_int64, _dp]),
    "mgo_ensemble_percentiles_values": (C.c_int32, [C.c_void_p, _dp, C.c_int32, C.c_int64, _dp, C.c_int32, _dp]),
    "mgo_ensemble_percentiles_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, _dp, C.c_int32,
                                                     C.c_void_p, C.c_void_p]),
    "mgo_sum_nodes_values": (C.c_int32, [C.c_void_p, _dp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp]),
    "mgo_serve_start": (C.c_int32, [C.c_void_p, C.POINTER(ServeCfg)]),
    "mgo_serve_submit": (C.c_int32, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "mgo_serve_update": (C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(DampingUpdate)]),
    "mgo_serve_poll": (C.c_int32, [C.c_void_p, C.POINTER(ResultInfo), C.c_int32, _ip]),
    "mgo_serve_get_stats": (C.c_int32, [C.c_void_p, C.POINTER(ServeStats)]),
    "mgo_serve_take": (C.c_int32, [C.c_void_p, C.c_int64, _dp, C.c_size_t]),
    "mgo_serve_updates_applied": (C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(DampingUpdate), C.c_int32, _ip]),
    "mgo_serve_stop": (C.c_int32, [C.c_void_p]),
    "mgo_clock_ms": (C.c_double, [C.c_void_p]),
}

_lib = None


def load(path=LIB_PATH):
    """Open the device library; raises if it was not built (no CPU fallback exists)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"memilio_amd: device library {path} is missing; run __graft_entry__.build() first")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def dptr(a):
    """double* of a C-contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "expected C-contiguous float64"
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "expected C-contiguous int32"
    return a.ctypes.data_as(_ip)


def integrator_cfg(kind=INTEGRATOR_CASH_KARP54, abs_tol=1e-10, rel_tol=1e-5, dt_min=DBL_MIN, dt_max=DBL_MAX):
    """Defaults of ControlledStepperWrapper (stepper_wrapper.h:83-91)."""
    return IntegratorCfg(kind, abs_tol, rel_tol, dt_min, dt_max)


def sim_cfg(t0=0.0, dt_graph=0.5, dt_node=0.1, horizon=0.0, integrator=None):
    return SimCfg(t0, dt_graph, dt_node, horizon, integrator if integrator is not None else integrator_cfg())
