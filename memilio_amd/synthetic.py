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


def _node_params(spec, k, rng, exposed_fraction=1e-3):
    """One model of the single-model workload: contacts U[0.5,10], stay times U[2,14], transmission
    probability U[0.02,0.2], group populations U[1e5,1e6], the given fraction initially exposed."""
    G = spec.G
    sc = spec.node_scalars[k]
    sc[capi.P_SEASONALITY] = 0.2
    sc[capi.P_START_DAY] = 0.0
    a = spec.age_params[k]
    for p, v in FIXED.items():
        a[p, :] = v
    for p in TIMES:
        a[p, :] = rng.uniform(2.0, 14.0, G)
    a[capi.A_TRANSMISSION_PROBABILITY_ON_CONTACT, :] = rng.uniform(0.02, 0.2, G)
    spec.contact_baseline[k] = rng.uniform(0.5, 10.0, (G, G))
    n = rng.uniform(1e5, 1e6, G)
    spec.populations[k, :, :] = 0.0
    spec.populations[k, :, E] = exposed_fraction * n
    spec.populations[k, :, S] = n - spec.populations[k, :, E]


def single_model(seed=0, G=6):
    """Plain single osecir model: G age groups, one node."""
    spec = GraphSpec(1, G)
    _node_params(spec, 0, np.random.default_rng(seed))
    return spec


def random_edges(rng, n, mean_degree=50.0, total_outflow=0.3):
    """Directed random graph: Poisson(mean_degree) out-degree (capped at n-1), targets without
    replacement, lognormal(-5, 1) commuter fractions rescaled so each node's outflow is <= total_outflow."""
    starts, ends, coeffs = [], [], []
    for s in range(n):
        d = min(int(rng.poisson(mean_degree)), n - 1)
        if d <= 0:
            continue
        others = np.delete(np.arange(n), s)
        tg = np.sort(rng.choice(others, size=d, replace=False))
        w = rng.lognormal(-5.0, 1.0, d)
        tot = w.sum()
        if tot > total_outflow:
            w = w * (total_outflow / tot)
        starts.append(np.full(d, s))
        ends.append(tg)
        coeffs.append(w)
    if not starts:
        return np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0)
    return (np.concatenate(starts).astype(np.int32), np.concatenate(ends).astype(np.int32),
            np.concatenate(coeffs))


def graph(num_nodes=400, G=6, seed=1, mean_degree=50.0):
    """Graph-ODE workload: num_nodes single-model nodes with their own seeds, random mobility edges."""
    spec = GraphSpec(num_nodes, G)
    for k in range(num_nodes):
        _node_params(spec, k, np.random.default_rng([seed, k]))
    s, e, c = random_edges(np.random.default_rng([seed, 10 ** 6]), num_nodes, mean_degree)
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
        for k in range(n):
            _node_params(spec, k, np.random.default_rng([seed, i, k]))
        if n > 1:
            s, e, c = random_edges(np.random.default_rng([seed, i, 10 ** 6]), n, min(50.0, n - 1))
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
