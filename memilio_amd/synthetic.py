"""Seeded synthetic workloads of the benchmark configurations (numpy PCG64, explicit seeds).

Parameters the workload does not draw are set to the values of the reference's graph example
(cpp/examples/ode_secir_graph.cpp:42-56): RelativeTransmissionNoSymptoms 0.7,
RecoveredPerInfectedNoSymptoms 0.09, RiskOfInfectionFromSymptomatic 0.25, MaxRisk 0.45,
SeverePerInfectedSymptoms 0.2, CriticalPerSevere 0.25, DeathsPerSevere 0, DeathsPerCritical 0.3,
Seasonality 0.2, StartDay 0; TestAndTraceCapacity and ICUCapacity keep their defaults (DBL_MAX).
"""
import numpy as np

from . import capi
from .graphspec import GraphSpec, S, E

FIXED = {
    capi.A_RELATIVE_TRANSMISSION_NO_SYMPTOMS: 0.7,
    capi.A_RECOVERED_PER_INFECTED_NO_SYMPTOMS: 0.09,
    capi.A_RISK_OF_INFECTION_FROM_SYMPTOMATIC: 0.25,
    capi.A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC: 0.45,
    capi.A_SEVERE_PER_INFECTED_SYMPTOMS: 0.2,
    capi.A_CRITICAL_PER_SEVERE: 0.25,
    capi.A_DEATHS_PER_SEVERE: 0.0,
    capi.A_DEATHS_PER_CRITICAL: 0.3,
}
TIMES = (capi.A_TIME_EXPOSED, capi.A_TIME_INFECTED_NO_SYMPTOMS, capi.A_TIME_INFECTED_SYMPTOMS,
         capi.A_TIME_INFECTED_SEVERE, capi.A_TIME_INFECTED_CRITICAL)


def _node_params(spec, rng, exposed_fraction=1e-3):
    """Every node of spec as one model of the single-model workload (each node its own draws):
    contacts U[0.5,10], stay times U[2,14], transmission probability U[0.02,0.2], group populations
    U[1e5,1e6], the given fraction initially exposed."""
    n, G = spec.num_nodes, spec.G
    spec.node_scalars[:, capi.P_SEASONALITY] = 0.2
    spec.node_scalars[:, capi.P_START_DAY] = 0.0
    a = spec.age_params
    for p, v in FIXED.items():
        a[:, p, :] = v
    for p in TIMES:
        a[:, p, :] = rng.uniform(2.0, 14.0, (n, G))
    a[:, capi.A_TRANSMISSION_PROBABILITY_ON_CONTACT, :] = rng.uniform(0.02, 0.2, (n, G))
    spec.contact_baseline[:] = rng.uniform(0.5, 10.0, (n, G, G))
    tot = rng.uniform(1e5, 1e6, (n, G))
    spec.populations[:] = 0.0
    spec.populations[:, :, E] = exposed_fraction * tot
    spec.populations[:, :, S] = tot - spec.populations[:, :, E]


def single_model(seed=0, G=6):
    """Plain single osecir model: G age groups, one node."""
    spec = GraphSpec(1, G)
    _node_params(spec, np.random.default_rng(seed))
    return spec


def random_edges(rng, n, mean_degree=50.0, total_outflow=0.3):
    """Directed random graph: Poisson(mean_degree) out-degree (capped at n-1), targets drawn without
    replacement among the other nodes, lognormal(-5, 1) commuter fractions rescaled so each node's
    outflow is <= total_outflow. Vectorised over all sources: draw with replacement, drop repeats,
    redraw the missing targets until every source has its degree."""
    deg = np.minimum(rng.poisson(mean_degree, n), n - 1).astype(np.int64)
    if n <= 1 or deg.sum() == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0)
    if n <= 128:
        # small graphs: a random order of the other nodes per source
        keys = rng.random((n, n))
        keys[np.arange(n), np.arange(n)] = np.inf
        order = np.argsort(keys, axis=1)
        src = np.repeat(np.arange(n), deg)
        pos = np.arange(src.size) - np.repeat(np.cumsum(deg) - deg, deg)
        pairs = np.unique(src * n + order[src, pos])
    else:
        pairs = np.zeros(0, np.int64)
        need = deg.copy()
        while need.sum() > 0:
            src = np.repeat(np.arange(n), need)
            t = rng.integers(0, n - 1, src.size)
            t = t + (t >= src)
            pairs = np.unique(np.concatenate([pairs, src * n + t]))
            have = np.bincount(pairs // n, minlength=n)
            need = np.maximum(deg - have, 0)
        # sources may have gained a few extra targets in the last round: keep the first deg per source
        s_all = pairs // n
        first = np.searchsorted(s_all, np.arange(n))
        rank = np.arange(pairs.size) - first[s_all]
        pairs = pairs[rank < deg[s_all]]
    starts = (pairs // n).astype(np.int32)
    ends = (pairs % n).astype(np.int32)
    w = rng.lognormal(-5.0, 1.0, pairs.size)
    tot = np.bincount(starts, weights=w, minlength=n)
    scale = np.where(tot > total_outflow, total_outflow / np.maximum(tot, 1e-300), 1.0)
    return starts, ends, w * scale[starts]


def graph(num_nodes=400, G=6, seed=1, mean_degree=50.0):
    """Graph-ODE workload: num_nodes single-model nodes with their own seeds, random mobility edges."""
    spec = GraphSpec(num_nodes, G)
    _node_params(spec, np.random.default_rng([seed, 0]))
    s, e, c = random_edges(np.random.default_rng([seed, 1]), num_nodes, mean_degree)
    spec.set_edges(s, e, c)
    return spec


def ensemble(base, num_samples=1024, seed=2):
    """Parameter study around a base graph: every node parameter scaled by U[0.8, 1.2], initial exposed
    fraction U[1e-4, 1e-2]; all samples share the base graph's edge arrays (one device copy)."""
    specs = []
    rng = np.random.default_rng(seed)
    n, G = base.num_nodes, base.G
    for _ in range(num_samples):
        sp = GraphSpec(n, G)
        sp.node_scalars = base.node_scalars.copy()
        f = rng.uniform(0.8, 1.2, base.age_params.shape)
        sp.age_params = base.age_params * f
        sp.contact_baseline = base.contact_baseline * rng.uniform(0.8, 1.2, base.contact_baseline.shape)
        tot = base.populations.sum(axis=2) * rng.uniform(0.8, 1.2, (n, G))
        frac = rng.uniform(1e-4, 1e-2, (n, G))
        sp.populations = np.zeros_like(base.populations)
        sp.populations[:, :, E] = frac * tot
        sp.populations[:, :, S] = tot - sp.populations[:, :, E]
        sp.edge_start, sp.edge_end, sp.edge_coeffs = base.edge_start, base.edge_end, base.edge_coeffs
        specs.append(sp)
    return specs


def uneven(num_instances=4096, seed=3, max_nodes=2000):
    """Independent graphs of uneven size: nodes log-uniform in [1, max_nodes], G uniform in 1..8,
    out-degree Poisson(min(50, n-1)), horizons U[30, 180] days (integer), rel tol in {1e-6, 1e-8, 1e-10}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(num_instances):
        n = int(np.exp(rng.uniform(0.0, np.log(max_nodes))))
        n = max(1, min(max_nodes, n))
        G = int(rng.integers(1, 9))
        spec = GraphSpec(n, G)
        _node_params(spec, np.random.default_rng([seed, i, 0]))
        if n > 1:
            s, e, c = random_edges(np.random.default_rng([seed, i, 1]), n, min(50.0, n - 1))
            spec.set_edges(s, e, c)
        horizon = float(rng.integers(30, 181))
        tol = float(rng.choice([1e-6, 1e-8, 1e-10]))
        out.append((spec, horizon, tol))
    return out


def ck_cfg(tol=1e-10, t0=0.0, dt_graph=0.5, dt_node=0.1):
    """Adaptive Cash-Karp (the reference's default core) with abs = rel = tol."""
    return capi.sim_cfg(t0=t0, dt_graph=dt_graph, dt_node=dt_node,
                        integrator=capi.integrator_cfg(capi.INTEGRATOR_CASH_KARP54, tol, tol))


def fixed_cfg(dt=0.1, t0=0.0, dt_graph=0.5):
    """Fixed-step Cash-Karp (ExplicitStepperWrapper<runge_kutta_cash_karp54>) with step dt."""
    return capi.sim_cfg(t0=t0, dt_graph=dt_graph, dt_node=dt,
                        integrator=capi.integrator_cfg(capi.INTEGRATOR_CASH_KARP54_FIXED, 0.0, 0.0, 0.0, 0.0))


def with_npis(spec, thresholds=((200.0, 0.3), (800.0, 0.6)), duration=14.0, delay=3.0, base_value=100000.0,
              directive_begin=None, directive_end=None, levels=(0, 1)):
    """Dynamic NPIs on every node of spec: per threshold (cases per base_value, damping value), one damping of
    that value with uniform group weights, at level levels[k % len(levels)]. Returns spec."""
    cfg = {"thresholds": [(thr, [(val, levels[k % len(levels)], 0, np.ones(spec.G))])
                          for k, (thr, val) in enumerate(thresholds)],
           "duration": duration, "delay": delay, "base_value": base_value}
    if directive_begin is not None:
        cfg["directive_begin"] = directive_begin
    if directive_end is not None:
        cfg["directive_end"] = directive_end
    spec.set_dynamic_npis([cfg] * spec.num_nodes)
    return spec


def with_edge_dampings(spec, dampings=((5.0, 0, 0, 0.3), (12.0, 1, 0, 0.5)), thresholds=((300.0, 0.4),),
                       duration=7.0, base_value=100000.0, minimum=0.2):
    """Mobility-coefficient dampings and edge dynamic NPIs on every out-edge (as osecir::draw_sample sets
    them): dampings (time, level, type, value for every group), NPIs of one damping per threshold, minimum
    coefficients = minimum * baseline. Returns spec."""
    G = spec.G
    spec.set_out_dampings([[(t, lev, typ, np.full(G, v)) for (t, lev, typ, v) in dampings]] * spec.num_nodes)
    if thresholds:
        spec.set_out_dynamic_npis([{"thresholds": [(thr, [(val, 0, 0, np.ones(G))]) for thr, val in thresholds],
                                    "duration": duration, "base_value": base_value}] * spec.num_nodes)
    spec.edge_coeff_minimum = minimum * np.asarray(spec.edge_coeffs, dtype=np.float64)
    spec._desc = None
    return spec
