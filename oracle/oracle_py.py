"""ctypes loader of the oracle (oracle/liboracle.so).

The oracle is the parity checker: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
import this module.
"""
import ctypes as C
import os

import numpy as np

from memilio_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("attempts", C.c_int64),
        ("rhs_evals", C.c_int64),
        ("negative_fixups", C.c_int64),
        ("returns", C.c_int64),
    ]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library {LIB_PATH} missing; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.orc_simulate.restype = C.c_int32
        L.orc_simulate.argtypes = [C.POINTER(capi.GraphDesc), C.POINTER(capi.SimCfg), C.c_double, C.c_int32, _dp, _dp]
        L.orc_simulate_daily.restype = C.c_int32
        L.orc_simulate_daily.argtypes = [C.POINTER(capi.GraphDesc), C.POINTER(capi.SimCfg), C.c_double, C.c_int32,
                                         _dp, C.POINTER(OrcStats)]
        L.orc_graph_run.restype = C.c_int32
        L.orc_graph_run.argtypes = [C.POINTER(capi.GraphDesc), C.POINTER(capi.SimCfg), C.c_double, C.c_int32, _dp,
                                    _ip, C.POINTER(OrcStats)]
        L.orc_graph_run_threads.restype = C.c_int32
        L.orc_graph_run_threads.argtypes = [C.POINTER(capi.GraphDesc), C.POINTER(capi.SimCfg), C.c_double, C.c_int32,
                                            C.c_int32, _dp, _ip, C.POINTER(OrcStats)]
        L.orc_graph_run_node_series.restype = C.c_int32
        L.orc_graph_run_node_series.argtypes = [C.POINTER(capi.GraphDesc), C.POINTER(capi.SimCfg), C.c_double,
                                                C.c_int32, C.c_int32, _dp, _dp]
        L.orc_ensemble_run.restype = C.c_int32
        L.orc_ensemble_run.argtypes = [C.POINTER(capi.EnsembleDesc), C.c_int32, _dp, C.POINTER(OrcStats)]
        L.orc_oseir_graph_series.restype = C.c_int32
        L.orc_oseir_graph_series.argtypes = [C.c_int32, C.c_int32, _dp, _dp, _dp, C.c_int32, _ip, _ip, _dp, C.c_double,
                                             C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, _dp, _dp]
        L.orc_edge_scenario.restype = C.c_int32
        L.orc_edge_scenario.argtypes = [C.POINTER(capi.GraphDesc), C.c_double, _dp, _dp]
        L.orc_integrate_test_problem.restype = C.c_int32
        L.orc_integrate_test_problem.argtypes = [C.c_int32, C.POINTER(capi.IntegratorCfg), C.c_double, C.c_double,
                                                 C.c_double, C.c_int32, _dp, _dp]
        L.orc_contact_at.restype = C.c_double
        L.orc_contact_at.argtypes = [C.POINTER(capi.GraphDesc), C.c_double, C.c_int32, C.c_int32]
        L.orc_implement_npis.restype = C.c_int32
        L.orc_implement_npis.argtypes = [C.c_int32, C.c_int32, C.c_int32, _ip, _dp, _ip, _ip, _dp, C.c_int32, _dp, _ip,
                                         _ip, _ip, _ip, C.c_int32, _dp, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                         _ip, _dp, _ip, _ip, _dp]
        L.orc_npi_frozen_run.restype = C.c_int32
        L.orc_npi_frozen_run.argtypes = [C.POINTER(capi.GraphDesc), C.c_double, C.c_double, C.c_int32, _dp, _dp, _ip]
        L.orc_edge_npi_calls.restype = C.c_int32
        L.orc_edge_npi_calls.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32, _dp,
                                         C.c_int32, _dp, _dp, _ip]
        _lib = L
    return _lib


_fma_lib = None


def fma_lib():
    """liboracle_fma.so: the same restatement compiled for x86-64-v3 with FMA contraction, a second valid
    compilation of the reference arithmetic (tests bound the drift between builds; bench.py times it as the
    optimised CPU build)."""
    global _fma_lib
    if _fma_lib is None:
        path = os.path.join(HERE, "liboracle_fma.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library {path} missing; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.orc_graph_run.restype = C.c_int32
        L.orc_graph_run.argtypes = [C.POINTER(capi.GraphDesc), C.POINTER(capi.SimCfg), C.c_double, C.c_int32, _dp,
                                    _ip, C.POINTER(OrcStats)]
        L.orc_ensemble_run.restype = C.c_int32
        L.orc_ensemble_run.argtypes = [C.POINTER(capi.EnsembleDesc), C.c_int32, _dp, C.POINTER(OrcStats)]
        L.orc_simulate_daily.restype = C.c_int32
        L.orc_simulate_daily.argtypes = [C.POINTER(capi.GraphDesc), C.POINTER(capi.SimCfg), C.c_double, C.c_int32,
                                         _dp, C.POINTER(OrcStats)]
        _fma_lib = L
    return _fma_lib


def graph_run_fma(spec, cfg, tmax):
    """graph_run with the FMA build (fma_lib)."""
    fma_lib()
    nd = spec.num_days(cfg.t0, tmax)
    out = np.zeros((nd, spec.num_nodes, spec.state_size()))
    got = C.c_int32(0)
    st = OrcStats()
    s = _fma_lib.orc_graph_run(C.byref(spec.desc()), C.byref(cfg), tmax, nd, capi.dptr(out), C.byref(got),
                               C.byref(st))
    if s != 0:
        raise RuntimeError(f"orc_graph_run (FMA build) failed: {s}")
    return out[:got.value], st


def edge_npi_calls(n, threshold, base, duration, value, weights, times, inf_rel):
    """Number of coefficient dampings after each apply_mobility NPI check (see orc_edge_npi_calls)."""
    w = np.ascontiguousarray(weights, dtype=np.float64)
    t = np.ascontiguousarray(times, dtype=np.float64)
    r = np.ascontiguousarray(inf_rel, dtype=np.float64)
    out = np.zeros(len(t), np.int32)
    st = lib().orc_edge_npi_calls(n, threshold, base, duration, value, len(w), capi.dptr(w), len(t), capi.dptr(t),
                                  capi.dptr(r), capi.iptr(out))
    if st != 0:
        raise RuntimeError(f"orc_edge_npi_calls: {st}")
    return out.tolist()


def implement_npis(rows, cols, lists, npis, begin, end, make_matrix=0, capacity=256):
    """implement_dynamic_npis on a group of damping lists.

    lists: per matrix, list of (time, level, type, coeffs[rows*cols]); npis: list of (value, level, type,
    matrix indices, weights). Returns the lists afterwards in the same form."""
    rc = rows * cols
    offs, times, levels, types, coeffs = [0], [], [], [], []
    for lst in lists:
        for (t, lev, typ, c) in lst:
            times.append(t)
            levels.append(lev)
            types.append(typ)
            coeffs.append(np.broadcast_to(np.asarray(c, np.float64), (rc,)))
        offs.append(len(times))
    nw = len(npis[0][4]) if npis else 0
    moffs, mats = [0], []
    for n in npis:
        mats += list(n[3])
        moffs.append(len(mats))
    i32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)       # noqa: E731
    f64 = lambda v: np.ascontiguousarray(v, dtype=np.float64)     # noqa: E731
    a = dict(offs=i32(offs), times=f64(times), levels=i32(levels), types=i32(types),
             coeffs=f64(np.stack(coeffs) if coeffs else np.zeros((0, rc))),
             nv=f64([n[0] for n in npis]), nl=i32([n[1] for n in npis]), nt=i32([n[2] for n in npis]),
             moffs=i32(moffs), mats=i32(mats), w=f64(np.stack([np.asarray(n[4], np.float64) for n in npis])))
    out_offs = np.zeros(len(lists) + 1, np.int32)
    ot, ol, oty = np.zeros(capacity), np.zeros(capacity, np.int32), np.zeros(capacity, np.int32)
    oc = np.zeros((capacity, rc))
    n = lib().orc_implement_npis(rows, cols, len(lists), capi.iptr(a["offs"]), capi.dptr(a["times"]),
                                 capi.iptr(a["levels"]), capi.iptr(a["types"]), capi.dptr(a["coeffs"]), len(npis),
                                 capi.dptr(a["nv"]), capi.iptr(a["nl"]), capi.iptr(a["nt"]), capi.iptr(a["moffs"]),
                                 capi.iptr(a["mats"]), nw, capi.dptr(a["w"]), begin, end, make_matrix, capacity,
                                 capi.iptr(out_offs), capi.dptr(ot), capi.iptr(ol), capi.iptr(oty), capi.dptr(oc))
    if n < 0 or n > capacity:
        raise RuntimeError(f"orc_implement_npis: {n}")
    return [[(float(ot[e]), int(ol[e]), int(oty[e]), oc[e].copy()) for e in range(out_offs[m], out_offs[m + 1])]
            for m in range(len(lists))]


def npi_frozen_run(spec, t0, tmax, query_times):
    """Simulation::advance(tmax) with dynamic NPIs on a constant state from t0: contact(0, 0) of node 0 at the
    query times afterwards and the number of dampings of its contact matrix."""
    q = np.ascontiguousarray(query_times, dtype=np.float64)
    out = np.zeros(len(q))
    nd = np.zeros(1, np.int32)
    st = lib().orc_npi_frozen_run(C.byref(spec.desc()), t0, tmax, len(q), capi.dptr(q), capi.dptr(out),
                                  capi.iptr(nd))
    if st != 0:
        raise RuntimeError(f"orc_npi_frozen_run: {st}")
    return out, int(nd[0])


def simulate(spec, cfg, tmax, capacity=200000):
    """Full TimeSeries (times[n], values[n][10G]) of node 0 of spec simulated alone."""
    n = spec.state_size()
    times = np.zeros(capacity)
    vals = np.zeros((capacity, n))
    np_ = lib().orc_simulate(C.byref(spec.desc()), C.byref(cfg), tmax, capacity, capi.dptr(times), capi.dptr(vals))
    if np_ < 0:
        raise RuntimeError(f"orc_simulate failed: {np_}")
    if np_ > capacity:
        return simulate(spec, cfg, tmax, np_)
    return times[:np_].copy(), vals[:np_].copy()


def simulate_daily(spec, cfg, tmax):
    """interpolate_simulation_result of orc_simulate: [day][10G] for node 0, and stats."""
    nd = spec.num_days(cfg.t0, tmax)
    out = np.zeros((nd, spec.state_size()))
    st = OrcStats()
    got = lib().orc_simulate_daily(C.byref(spec.desc()), C.byref(cfg), tmax, nd, capi.dptr(out), C.byref(st))
    if got < 0:
        raise RuntimeError(f"orc_simulate_daily failed: {got}")
    return out[:got], st


def graph_run(spec, cfg, tmax):
    """Daily values [day][node][10G] and stats of make_mobility_sim(...).advance(tmax)."""
    nd = spec.num_days(cfg.t0, tmax)
    out = np.zeros((nd, spec.num_nodes, spec.state_size()))
    got = C.c_int32(0)
    st = OrcStats()
    s = lib().orc_graph_run(C.byref(spec.desc()), C.byref(cfg), tmax, nd, capi.dptr(out), C.byref(got), C.byref(st))
    if s != 0:
        raise RuntimeError(f"orc_graph_run failed: {s}")
    return out[:got.value], st


def graph_run_threads(spec, cfg, tmax, node_threads, fma=False):
    """graph_run with the node phase on node_threads OpenMP threads (bitwise equal to graph_run); fma: the FMA
    build (fma_lib)."""
    L = fma_lib() if fma else lib()
    if fma and not getattr(L, "_threads_bound", False):
        L.orc_graph_run_threads.restype = C.c_int32
        L.orc_graph_run_threads.argtypes = [C.POINTER(capi.GraphDesc), C.POINTER(capi.SimCfg), C.c_double,
                                            C.c_int32, C.c_int32, _dp, _ip, C.POINTER(OrcStats)]
        L._threads_bound = True
    nd = spec.num_days(cfg.t0, tmax)
    out = np.zeros((nd, spec.num_nodes, spec.state_size()))
    got = C.c_int32(0)
    st = OrcStats()
    s = L.orc_graph_run_threads(C.byref(spec.desc()), C.byref(cfg), tmax, node_threads, nd, capi.dptr(out),
                                C.byref(got), C.byref(st))
    if s != 0:
        raise RuntimeError(f"orc_graph_run_threads failed: {s}")
    return out[:got.value], st


def graph_node_series(spec, cfg, tmax, node, capacity=400000):
    n = spec.state_size()
    times = np.zeros(capacity)
    vals = np.zeros((capacity, n))
    np_ = lib().orc_graph_run_node_series(C.byref(spec.desc()), C.byref(cfg), tmax, node, capacity,
                                          capi.dptr(times), capi.dptr(vals))
    if np_ < 0:
        raise RuntimeError(f"orc_graph_run_node_series failed: {np_}")
    if np_ > capacity:
        return graph_node_series(spec, cfg, tmax, node, np_)
    return times[:np_].copy(), vals[:np_].copy()


def ensemble_run(specs, cfgs, tmaxs, threads=0, want_output=True, fma=False):
    """ParameterStudy over pre-sampled graphs; returns (list of [day][node][10G] arrays or None, stats).
    threads: OpenMP threads over runs (0 = all); fma: run the FMA build (fma_lib)."""
    nr = len(specs)
    graphs = (capi.GraphDesc * nr)(*[s.desc() for s in specs])
    cfg_arr = (capi.SimCfg * nr)(*cfgs)
    tm = np.ascontiguousarray(tmaxs, dtype=np.float64)
    desc = capi.EnsembleDesc(nr, graphs, cfg_arr, capi.dptr(tm))
    sizes = [s.num_days(c.t0, t) * s.num_nodes * s.state_size() for s, c, t in zip(specs, cfgs, tm)]
    out = np.zeros(sum(sizes)) if want_output else None
    st = OrcStats()
    L = fma_lib() if fma else lib()
    s = L.orc_ensemble_run(C.byref(desc), threads, capi.dptr(out) if want_output else None, C.byref(st))
    if s != 0:
        raise RuntimeError(f"orc_ensemble_run failed: {s}")
    if not want_output:
        return None, st
    res, o = [], 0
    for spec, c, t, sz in zip(specs, cfgs, tm, sizes):
        res.append(out[o:o + sz].reshape(spec.num_days(c.t0, t), spec.num_nodes, spec.state_size()))
        o += sz
    return res, st


def oseir_graph_series(params, contact, populations, edge_start, edge_end, edge_coeffs, t0, dt_graph, dt_node, tmax,
                       node, capacity=20000):
    params = np.ascontiguousarray(params, dtype=np.float64)
    contact = np.ascontiguousarray(contact, dtype=np.float64)
    populations = np.ascontiguousarray(populations, dtype=np.float64)
    es = np.ascontiguousarray(edge_start, dtype=np.int32)
    ee = np.ascontiguousarray(edge_end, dtype=np.int32)
    ec = np.ascontiguousarray(edge_coeffs, dtype=np.float64)
    nn, G = populations.shape[0], populations.shape[1]
    times = np.zeros(capacity)
    vals = np.zeros((capacity, 4 * G))
    np_ = lib().orc_oseir_graph_series(nn, G, capi.dptr(params), capi.dptr(contact), capi.dptr(populations),
                                       len(es), capi.iptr(es), capi.iptr(ee), capi.dptr(ec), t0, dt_graph, dt_node,
                                       tmax, node, capacity, capi.dptr(times), capi.dptr(vals))
    if np_ < 0:
        raise RuntimeError("orc_oseir_graph_series failed")
    return times[:np_].copy(), vals[:np_].copy()


def edge_scenario(spec, t, coeffs):
    n = spec.state_size()
    out = np.zeros((6, n))
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    s = lib().orc_edge_scenario(C.byref(spec.desc()), t, capi.dptr(c), capi.dptr(out))
    if s != 0:
        raise RuntimeError("orc_edge_scenario failed")
    return out.reshape(3, 2, n)


def integrate_test_problem(cfg, t0, tmax, dt, capacity=100000):
    times = np.zeros(capacity)
    vals = np.zeros(capacity)
    np_ = lib().orc_integrate_test_problem(0, C.byref(cfg), t0, tmax, dt, capacity, capi.dptr(times), capi.dptr(vals))
    if np_ < 0:
        raise RuntimeError("orc_integrate_test_problem failed")
    return times[:np_].copy(), vals[:np_].copy()


def contact_at(spec, t, i, j):
    return lib().orc_contact_at(C.byref(spec.desc()), t, i, j)


# ------------------------------------------------------------------------------- ensemble analysis
# cpp/memilio/data/analyze_result.h:267-343, restated over arrays results[run][day][node][element]

def ensemble_mean(results):
    """ensemble_mean (analyze_result.h:289-315): per position, sum over runs in run order of value/num_runs."""
    n = len(results)
    acc = np.zeros_like(results[0])
    for r in range(n):
        acc = acc + results[r] / float(n)
    return acc


def ensemble_percentile(results, p):
    """ensemble_percentile (analyze_result.h:316-343): sorted run values at index size_t(num_runs * p)."""
    n = len(results)
    s = np.sort(np.stack(results), axis=0)
    return s[int(float(n) * p)]


def sum_nodes(results):
    """sum_nodes (analyze_result.h:267-287): per run and day, nodes summed in node order: [run][day][element]."""
    out = []
    for res in results:
        acc = np.zeros((res.shape[0], res.shape[2]))
        for v in range(res.shape[1]):
            acc = acc + res[:, v, :]
        out.append(acc)
    return np.stack(out)
