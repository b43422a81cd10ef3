"""Python mirror of `memilio.simulation.osecir` for the Graph-ODE path, backed by the HIP engine.

Reference bindings mirrored (pycode/memilio-simulation/memilio/simulation/bindings/models/osecir.cpp,
bindings/mobility/graph_simulation.h:35-52, bindings/compartments/parameter_studies.h:92-124):
InfectionState, Model (parameters by name, populations indexed by (AgeGroup, InfectionState),
apply_constraints), ModelGraph, MobilityGraph, MobilitySimulation(graph, t0, dt).advance(tmax),
GraphParameterStudy(graph, t0, tmax, dt, num_runs).run([handle]), simulate(t0, tmax, dt, model),
interpolate_simulation_result, write_graph / read_graph (parameter-graph JSON files).

Differences a caller sees: results are daily (the values interpolate_simulation_result would give for the
reference's step-wise series); GraphParameterStudy.run takes an optional `sample(graph, run_idx)` callback
in place of the reference's draw_sample over parameter distributions, and runs every sampled graph in one
device batch; a MobilitySimulation advances up to the horizon of its first advance unless `horizon` is
given. Every solve runs on the device; a missing engine library raises.
"""
import copy
import enum

import numpy as np

from .. import capi
from ..engine import Context
from ..graphspec import GraphSpec, NC
from . import (AgeGroup, ContactMatrix, ContactMatrixGroup, DampingSampling, DynamicNPIs, MobilityParameters,
               ParameterDistributionConstant, ParameterDistributionNormal, TimeSeries, UncertainContactMatrix,
               UncertainValue, apply_dampings, make_mobility_damping_vector)
from . import interpolate_simulation_result as _interp_ts

__all__ = ["InfectionState", "Model", "ModelGraph", "MobilityGraph", "MobilitySimulation", "GraphParameterStudy",
           "ParameterStudy", "Simulation", "simulate", "interpolate_simulation_result", "interpolate_ensemble_results",
           "ensemble_mean", "ensemble_percentile", "ControlledStepperWrapper",
           "RKIntegratorCore", "ExplicitStepperWrapper", "EulerIntegratorCore", "write_graph", "read_graph",
           "draw_sample", "set_params_distributions_normal", "MultiIndex_PopulationsArray"]


class InfectionState(enum.IntEnum):
    Susceptible = 0
    Exposed = 1
    InfectedNoSymptoms = 2
    InfectedNoSymptomsConfirmed = 3
    InfectedSymptoms = 4
    InfectedSymptomsConfirmed = 5
    InfectedSevere = 6
    InfectedCritical = 7
    Recovered = 8
    Dead = 9
    Count = 10


_SCALARS = {
    "StartDay": capi.P_START_DAY,
    "Seasonality": capi.P_SEASONALITY,
    "ICUCapacity": capi.P_ICU_CAPACITY,
    "TestAndTraceCapacity": capi.P_TEST_AND_TRACE_CAPACITY,
    "TestAndTraceCapacityMaxRisk": capi.P_TEST_AND_TRACE_MAX_RISK,
}
_AGES = {
    "TimeExposed": capi.A_TIME_EXPOSED,
    "TimeInfectedNoSymptoms": capi.A_TIME_INFECTED_NO_SYMPTOMS,
    "TimeInfectedSymptoms": capi.A_TIME_INFECTED_SYMPTOMS,
    "TimeInfectedSevere": capi.A_TIME_INFECTED_SEVERE,
    "TimeInfectedCritical": capi.A_TIME_INFECTED_CRITICAL,
    "TransmissionProbabilityOnContact": capi.A_TRANSMISSION_PROBABILITY_ON_CONTACT,
    "RelativeTransmissionNoSymptoms": capi.A_RELATIVE_TRANSMISSION_NO_SYMPTOMS,
    "RecoveredPerInfectedNoSymptoms": capi.A_RECOVERED_PER_INFECTED_NO_SYMPTOMS,
    "RiskOfInfectionFromSymptomatic": capi.A_RISK_OF_INFECTION_FROM_SYMPTOMATIC,
    "MaxRiskOfInfectionFromSymptomatic": capi.A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC,
    "SeverePerInfectedSymptoms": capi.A_SEVERE_PER_INFECTED_SYMPTOMS,
    "CriticalPerSevere": capi.A_CRITICAL_PER_SEVERE,
    "DeathsPerSevere": capi.A_DEATHS_PER_SEVERE,
    "DeathsPerCritical": capi.A_DEATHS_PER_CRITICAL,
}


# ----------------------------------------------------------------------------------------- integrators
class ControlledStepperWrapper:
    """Adaptive Cash-Karp 5(4) with odeint's controller: the reference's default core."""
    kind = capi.INTEGRATOR_CASH_KARP54

    def __init__(self, abs_tol=1e-10, rel_tol=1e-5, dt_min=capi.DBL_MIN, dt_max=capi.DBL_MAX):
        self.abs_tol, self.rel_tol, self.dt_min, self.dt_max = abs_tol, rel_tol, dt_min, dt_max

    def cfg(self):
        return capi.integrator_cfg(self.kind, self.abs_tol, self.rel_tol, self.dt_min, self.dt_max)


class RKIntegratorCore(ControlledStepperWrapper):
    """RKIntegratorCore with the default Runge-Kutta-Fehlberg 4(5) tableau."""
    kind = capi.INTEGRATOR_RKF45


class ExplicitStepperWrapper:
    """Fixed-step Cash-Karp (the simulation's dt)."""

    def cfg(self):
        return capi.integrator_cfg(capi.INTEGRATOR_CASH_KARP54_FIXED, 0.0, 0.0, 0.0, 0.0)


class EulerIntegratorCore:
    def cfg(self):
        return capi.integrator_cfg(capi.INTEGRATOR_EULER, 0.0, 0.0, 0.0, 0.0)


_ctx = None


def _context():
    """Process-wide device context (device 0), created on first use."""
    global _ctx
    if _ctx is None:
        _ctx = Context()
    return _ctx


# ----------------------------------------------------------------------------------------- model
class _AgeParameter:
    """Per-age-group values of one parameter: p[AgeGroup(i)] = v, p[AgeGroup(i)].value, and the groups'
    distributions (p[AgeGroup(i)].set_distribution(...))."""

    def __init__(self, arr, dists):
        self._a = arr
        self._d = dists   # {group: distribution}

    def __getitem__(self, g):
        i = int(g.get() if isinstance(g, AgeGroup) else g)
        return UncertainValue(lambda: float(self._a[i]), lambda v: self._a.__setitem__(i, v),
                              lambda: self._d.get(i), lambda d: self._d.__setitem__(i, d))

    def __setitem__(self, g, v):
        i = int(g.get() if isinstance(g, AgeGroup) else g)
        self._a[i] = float(v.value if isinstance(v, UncertainValue) else v)

    def get_num_elements(self):
        return len(self._a)

    def numel(self):
        return len(self._a)


class Parameters:
    """osecir::Parameters by name (parameters.h:46-394, 400-450)."""

    def __init__(self, num_groups):
        object.__setattr__(self, "_G", int(num_groups))
        from ..graphspec import default_age_params, default_scalars
        object.__setattr__(self, "_sc", default_scalars())
        object.__setattr__(self, "_age", default_age_params(self._G))
        object.__setattr__(self, "ContactPatterns",
                           UncertainContactMatrix(ContactMatrixGroup(1, self._G)))
        object.__setattr__(self, "DynamicNPIsInfectedSymptoms", DynamicNPIs())
        object.__setattr__(self, "_dist", {})   # distributions: name -> dist (scalars), name -> {group: dist}

    def __getattr__(self, name):
        if name in _SCALARS:
            k = _SCALARS[name]
            if name == "StartDay":   # a plain number in the reference (parameters.h:47-62)
                return UncertainValue(lambda: float(self._sc[k]), lambda v: self._sc.__setitem__(k, v))
            return UncertainValue(lambda: float(self._sc[k]), lambda v: self._sc.__setitem__(k, v),
                                  lambda: self._dist.get(name), lambda d: self._dist.__setitem__(name, d))
        if name in _AGES:
            return _AgeParameter(self._age[_AGES[name]], self._dist.setdefault(name, {}))
        raise AttributeError(f"osecir parameter {name!r} is not on the device path")

    def __setattr__(self, name, v):
        if name in _SCALARS:
            self._sc[_SCALARS[name]] = float(v.value if isinstance(v, UncertainValue) else v)
        elif name in _AGES:
            self._age[_AGES[name]][:] = np.asarray(v, dtype=np.float64)
        elif name == "ContactPatterns":
            object.__setattr__(self, name, v if isinstance(v, UncertainContactMatrix) else UncertainContactMatrix(v))
        elif name == "DynamicNPIsInfectedSymptoms":
            if not isinstance(v, DynamicNPIs):
                raise TypeError("DynamicNPIsInfectedSymptoms takes a DynamicNPIs")
            object.__setattr__(self, name, v)
        else:
            raise AttributeError(f"osecir parameter {name!r} is not on the device path")

    def get_num_groups(self):
        return self._G

    # commuter detection (parameters.h:417-450)
    def get_commuter_nondetection(self):
        return float(self._sc[capi.P_COMMUTER_NONDETECTION])

    def set_commuter_nondetection(self, v):
        self._sc[capi.P_COMMUTER_NONDETECTION] = float(v)

    def get_start_commuter_detection(self):
        return float(self._sc[capi.P_START_COMMUTER_DETECTION])

    def set_start_commuter_detection(self, v):
        self._sc[capi.P_START_COMMUTER_DETECTION] = float(v)

    def get_end_commuter_detection(self):
        return float(self._sc[capi.P_END_COMMUTER_DETECTION])

    def set_end_commuter_detection(self, v):
        self._sc[capi.P_END_COMMUTER_DETECTION] = float(v)

    def __deepcopy__(self, memo):
        p = Parameters(self._G)
        object.__setattr__(p, "_sc", self._sc.copy())
        object.__setattr__(p, "_age", self._age.copy())
        object.__setattr__(p, "ContactPatterns", copy.deepcopy(self.ContactPatterns, memo))
        object.__setattr__(p, "DynamicNPIsInfectedSymptoms", copy.deepcopy(self.DynamicNPIsInfectedSymptoms, memo))
        object.__setattr__(p, "_dist", copy.deepcopy(self._dist, memo))
        return p

    def apply_constraints(self):
        """Clamp parameters to admissible values as osecir::Parameters::apply_constraints
        (parameters.h:465-612): seasonality outside [0, 0.5] -> 0, negative capacities -> 0, stay times
        below 0.1 -> 0.1, RelativeTransmissionNoSymptoms < 0 -> 0, the other probabilities outside [0, 1]
        -> 0 (MaxRiskOfInfectionFromSymptomatic is not checked), DeathsPerSevere -> 0 where
        CriticalPerSevere + DeathsPerSevere > 1. Returns True if anything was corrected."""
        corrected = False
        sc, a = self._sc, self._age
        if not 0.0 <= sc[capi.P_SEASONALITY] <= 0.5:
            sc[capi.P_SEASONALITY] = 0.0
            corrected = True
        for k in (capi.P_ICU_CAPACITY, capi.P_TEST_AND_TRACE_CAPACITY):
            if sc[k] < 0.0:
                sc[k] = 0.0
                corrected = True
        if sc[capi.P_TEST_AND_TRACE_MAX_RISK] < 0.0:
            sc[capi.P_TEST_AND_TRACE_MAX_RISK] = 0.0
            corrected = True
        for k in (capi.A_TIME_EXPOSED, capi.A_TIME_INFECTED_NO_SYMPTOMS, capi.A_TIME_INFECTED_SYMPTOMS,
                  capi.A_TIME_INFECTED_SEVERE, capi.A_TIME_INFECTED_CRITICAL):
            bad = a[k] < 0.1
            if bad.any():
                a[k][bad] = 0.1
                corrected = True
        bad = a[capi.A_RELATIVE_TRANSMISSION_NO_SYMPTOMS] < 0.0
        if bad.any():
            a[capi.A_RELATIVE_TRANSMISSION_NO_SYMPTOMS][bad] = 0.0
            corrected = True
        for k in (capi.A_TRANSMISSION_PROBABILITY_ON_CONTACT, capi.A_RECOVERED_PER_INFECTED_NO_SYMPTOMS,
                  capi.A_RISK_OF_INFECTION_FROM_SYMPTOMATIC, capi.A_SEVERE_PER_INFECTED_SYMPTOMS,
                  capi.A_CRITICAL_PER_SEVERE, capi.A_DEATHS_PER_SEVERE, capi.A_DEATHS_PER_CRITICAL):
            bad = (a[k] < 0.0) | (a[k] > 1.0)
            if bad.any():
                a[k][bad] = 0.0
                corrected = True
        bad = a[capi.A_CRITICAL_PER_SEVERE] + a[capi.A_DEATHS_PER_SEVERE] > 1.0
        if bad.any():
            a[capi.A_DEATHS_PER_SEVERE][bad] = 0.0
            corrected = True
        return corrected


class Populations:
    """Compartment values indexed by (AgeGroup, InfectionState); flat index g * 10 + state."""

    def __init__(self, num_groups):
        self._G = int(num_groups)
        self._v = np.zeros((self._G, NC))
        self._dist = {}   # distributions of the compartments (draw_sample, graph files), by flat index

    @staticmethod
    def _idx(key):
        g, s = key
        return int(g.get() if isinstance(g, AgeGroup) else g), int(s)

    def __getitem__(self, key):
        g, s = self._idx(key)
        f = g * NC + s
        return UncertainValue(lambda: float(self._v[g, s]), lambda v: self._v.__setitem__((g, s), v),
                              lambda: self._dist.get(f), lambda d: self._dist.__setitem__(f, d))

    def __setitem__(self, key, v):
        g, s = self._idx(key)
        self._v[g, s] = float(v.value if isinstance(v, UncertainValue) else v)

    def numel(self):
        return self._v.size

    def get_num_compartments(self):
        return self._v.size

    def get_flat_index(self, key):
        g, s = self._idx(key)
        return g * NC + s

    def get_total(self):
        return float(self._v.sum())

    def get_group_total_AgeGroup(self, g):
        return float(self._v[int(g.get() if isinstance(g, AgeGroup) else g)].sum())

    def set_difference_from_total(self, key, total):
        self[key] = 0.0
        self[key] = total - self.get_total()

    def set_difference_from_group_total_AgeGroup(self, key, total):
        g, _ = self._idx(key)
        self[key] = 0.0
        self[key] = total - float(self._v[g].sum())

    def get_compartments(self):
        return self._v.ravel().copy()


class Model:
    """osecir::Model: parameters and initial populations."""

    def __init__(self, num_agegroups):
        self.parameters = Parameters(num_agegroups)
        self.populations = Populations(num_agegroups)

    def apply_constraints(self):
        return self.parameters.apply_constraints()

    def __deepcopy__(self, memo):
        m = Model.__new__(Model)
        m.parameters = copy.deepcopy(self.parameters, memo)
        m.populations = copy.deepcopy(self.populations, memo)
        return m

    def check_constraints(self):
        test = copy.deepcopy(self)
        return test.apply_constraints()

    @property
    def num_groups(self):
        return self.parameters.get_num_groups()


# ----------------------------------------------------------------------------------------- graph
class _Node:
    def __init__(self, id, prop):
        self.id = id
        self.property = prop


class _Edge:
    def __init__(self, start, end, prop):
        self.start_node_idx = start
        self.end_node_idx = end
        self.property = prop


class _Graph:
    def __init__(self):
        self._nodes, self._edges = [], []

    @property
    def num_nodes(self):
        return len(self._nodes)

    @property
    def num_edges(self):
        return len(self._edges)

    def get_node(self, i):
        return self._nodes[i]

    def get_edge(self, i):
        return self._edges[i]

    def get_num_out_edges(self, i):
        return sum(1 for e in self._edges if e.start_node_idx == i)

    def get_out_edge(self, i, k):
        return [e for e in self._edges if e.start_node_idx == i][k]

    def add_edge(self, start, end, coeffs):
        """Edges stay sorted by (start, end); an existing edge is replaced (graph.h:231-241)."""
        if not (0 <= start < self.num_nodes and 0 <= end < self.num_nodes):
            raise IndexError("add_edge: node index out of range")
        # the graph keeps its own copy, as Graph::add_edge copies the edge property (graph.h:231-241)
        prop = copy.deepcopy(coeffs) if isinstance(coeffs, MobilityParameters) else MobilityParameters(coeffs)
        for k, e in enumerate(self._edges):
            if (e.start_node_idx, e.end_node_idx) == (start, end):
                self._edges[k] = _Edge(start, end, prop)
                return
            if (e.start_node_idx, e.end_node_idx) > (start, end):
                self._edges.insert(k, _Edge(start, end, prop))
                return
        self._edges.append(_Edge(start, end, prop))


class ModelGraph(_Graph):
    """Graph<osecir::Model, MobilityParameters>: the parameter graph of a study."""

    def add_node(self, id, model):
        self._nodes.append(_Node(id, copy.deepcopy(model)))


class SimulationNode:
    """SimulationNode<osecir::Simulation>: the model, its initial time and step, and its result."""

    def __init__(self, model, t0=0.0, dt=0.1):
        self.model = copy.deepcopy(model)
        self.t0 = float(t0)
        self.dt = float(dt)
        self.result = TimeSeries(model.populations.numel())
        self.result.add_time_point(self.t0, model.populations.get_compartments())


class MobilityGraph(_Graph):
    """Graph<SimulationNode<osecir::Simulation>, MobilityEdge>."""

    def add_node(self, id, model, t0=0.0, dt=0.1):
        self._nodes.append(_Node(id, SimulationNode(model, t0, dt)))


def _spec_from(models, edges):
    """GraphSpec (mgo_graph_desc arrays) of a list of models and (start, end, coeffs) edges."""
    G = models[0].num_groups
    if any(m.num_groups != G for m in models):
        raise ValueError("all nodes of a graph need the same number of age groups")
    spec = GraphSpec(len(models), G)
    per_node_damp = []
    any_damp = False
    minima = np.zeros((len(models), G, G))
    for v, m in enumerate(models):
        p = m.parameters
        spec.node_scalars[v] = p._sc
        spec.age_params[v] = p._age
        spec.populations[v] = m.populations._v
        grp = p.ContactPatterns.cont_freq_mat
        damped = [k for k in range(grp.num_matrices) if grp[k].dampings]
        dn = p.DynamicNPIsInfectedSymptoms
        targets = {k for _, ds in dn.threshold for d in ds for k in d.matrix_indices}
        if any(len(d.matrix_indices) != 1 for _, ds in dn.threshold for d in ds) or len(targets) > 1 or \
                any(k >= grp.num_matrices for k in targets):
            raise ValueError("dynamic NPIs of a node must damp one and the same contact matrix on the device")
        damped = sorted(set(damped) | targets)
        if len(damped) > 1:
            raise ValueError("more than one damped contact matrix in a node is not supported on the device")
        B = sum(grp[k].baseline for k in range(grp.num_matrices))
        if damped:
            k0 = damped[0]
            M = grp[k0].minimum + sum(grp[k].baseline for k in range(grp.num_matrices) if k != k0)
        else:
            M = sum(grp[k].minimum for k in range(grp.num_matrices))
        spec.contact_baseline[v] = B
        minima[v] = M
        lst = []
        if damped:
            for d in grp[damped[0]].dampings:
                c = np.broadcast_to(d.coeffs, (G, G)) if d.coeffs.size == 1 else d.coeffs
                lst.append((d.time, d.level, d.type, c))
            any_damp = True
        per_node_damp.append(lst)
    if minima.any():
        spec.contact_minimum = minima
    if any_damp:
        spec.set_dampings(per_node_damp)
    npi_cfgs = [_npi_cfg(m.parameters.DynamicNPIsInfectedSymptoms) for m in models]
    if any(c["thresholds"] for c in npi_cfgs):
        spec.set_dynamic_npis(npi_cfgs)
    if edges:
        _edges_into(spec, edges, [m.populations for m in models])
    return spec


def _npi_cfg(dn):
    return {"thresholds": [(thr, [(float(d.value.value), d.level, d.type, d.group_weights) for d in ds])
                           for thr, ds in dn.threshold],
            "duration": dn.duration, "delay": dn.implementation_delay, "base_value": dn.base_value,
            "directive_begin": dn.directive_begin, "directive_end": dn.directive_end}


def _edges_into(spec, edges, populations):
    """Edges (start, end, MobilityParameters or coefficient vector) into the descriptor: baseline B and minimum
    M of the folded coefficient group, and per source node the out-edge dampings (per age group) and edge
    NPIs, which every out-edge of a node must share on the device."""
    G, n = spec.G, spec.num_nodes
    s = np.array([e[0] for e in edges], np.int32)
    d = np.array([e[1] for e in edges], np.int32)
    B, M, out = [], [], [None] * n
    for (start, _, prop) in edges:
        if not isinstance(prop, MobilityParameters):
            B.append(np.broadcast_to(np.asarray(prop, np.float64), (NC * G,)))
            M.append(np.zeros(NC * G))
            key = ((), None)
        else:
            grp = prop.coefficients
            damped = [k for k in range(grp.num_matrices) if grp[k].dampings]
            if len(damped) > 1:
                raise ValueError("more than one damped mobility coefficient vector in an edge is not supported")
            b = sum(np.broadcast_to(grp[k].baseline, (NC * G,)) for k in range(grp.num_matrices))
            if damped:
                k0 = damped[0]
                m = grp[k0].minimum + sum(grp[k].baseline for k in range(grp.num_matrices) if k != k0)
            else:
                m = sum(grp[k].minimum for k in range(grp.num_matrices))
            B.append(b)
            M.append(np.broadcast_to(m, (NC * G,)))
            lst = []
            for dm in (grp[damped[0]].dampings if damped else []):
                v = np.broadcast_to(dm.coeffs, (NC * G,)).reshape(G, NC)
                if not np.all(v == v[:, :1]):
                    raise ValueError("mobility dampings must be uniform within an age group on the device")
                lst.append((dm.time, dm.level, dm.type, tuple(v[:, 0])))
            dn = prop.dynamic_npis_infected
            npis = _npi_cfg(dn) if dn.threshold else None
            key = (tuple(lst), repr(npis))
        if out[start] is None:
            out[start] = (key, prop)
        elif out[start][0] != key:
            raise ValueError(f"out-edges of node {start} differ in their dampings or dynamic NPIs; the device "
                             "shares one list per node")
    spec.set_edges(s, d, np.stack(B))
    Ms = np.stack(M)
    if Ms.any():
        spec.edge_coeff_minimum = Ms
    if any(o is not None and o[0][0] for o in out):
        spec.set_out_dampings([[(t, lev, typ, np.asarray(v)) for (t, lev, typ, v) in o[0][0]] if o else []
                               for o in out])
    if any(o is not None and isinstance(o[1], MobilityParameters) and o[1].dynamic_npis_infected.threshold
           for o in out):
        spec.set_out_dynamic_npis([_npi_cfg(o[1].dynamic_npis_infected)
                                   if o is not None and isinstance(o[1], MobilityParameters) else None for o in out])
    spec._desc = None


def _integrator_of(graph, integrator):
    if integrator is not None:
        return integrator.cfg()
    return ControlledStepperWrapper().cfg()


def _fill_results(graph, daily, day0, t_end=None, state=None):
    for v in range(graph.num_nodes):
        node = graph.get_node(v).property
        ts = TimeSeries(node.result.get_num_elements())
        for k in range(daily.shape[0]):
            ts.add_time_point(day0 + k, daily[k, v])
        if state is not None and (daily.shape[0] == 0 or abs(day0 + daily.shape[0] - 1 - t_end) > 1e-10):
            ts.add_time_point(t_end, state[v])
        node.result = ts


class MobilitySimulation:
    """make_mobility_sim(t0, dt, graph): GraphSimulation with daily exchange for dt = 0.5."""

    def __init__(self, graph, t0=0.0, dt=1.0, integrator=None, horizon=None):
        self.graph = graph
        self.t = float(t0)
        self.t0 = float(t0)
        self.dt = float(dt)
        self._integrator = integrator
        self._horizon = horizon
        self._sim = None

    def advance(self, tmax):
        from ..engine import GraphSim
        if self._sim is None:
            nodes = [self.graph.get_node(v).property for v in range(self.graph.num_nodes)]
            dts = {n.dt for n in nodes}
            if len(dts) != 1:
                raise ValueError("nodes of one graph need the same initial step on the device")
            edges = [(e.start_node_idx, e.end_node_idx, e.property) for e in self.graph._edges]
            self._spec = _spec_from([n.model for n in nodes], edges)
            cfg = capi.sim_cfg(t0=self.t0, dt_graph=self.dt, dt_node=dts.pop(),
                               integrator=_integrator_of(self.graph, self._integrator))
            hz = max(float(tmax), float(self._horizon or 0.0))
            self._sim = GraphSim(_context(), self._spec, cfg, horizon=hz)
        self._sim.advance(float(tmax))
        self.t = self._sim.get_t()
        _fill_results(self.graph, self._sim.daily(), int(np.ceil(self.t0 - 1e-10)), self.t, self._sim.state())


class Simulation:
    """osecir::Simulation of one model; advance(tmax) runs t0 -> tmax on the device (once)."""

    def __init__(self, model, t0=0.0, dt=0.1, integrator=None):
        self.model = copy.deepcopy(model)
        self.t0, self.dt = float(t0), float(dt)
        self.integrator = integrator
        self.result = TimeSeries(model.populations.numel())
        self.result.add_time_point(self.t0, model.populations.get_compartments())

    def advance(self, tmax):
        spec = _spec_from([self.model], [])
        cfg = capi.sim_cfg(t0=self.t0, dt_node=self.dt, horizon=float(tmax),
                           integrator=(self.integrator or ControlledStepperWrapper()).cfg())
        daily = _context().simulate_single(spec, cfg, float(tmax))[:, 0, :]
        d0 = int(np.ceil(self.t0 - 1e-10))
        self.result = TimeSeries(self.model.populations.numel(), [d0 + k for k in range(daily.shape[0])], daily)
        return self.result


def simulate(t0, tmax, dt, model, integrator=None):
    """osecir::simulate(t0, tmax, dt, model[, core]) -> daily TimeSeries (model.h:385-390)."""
    return Simulation(model, t0, dt, integrator).advance(tmax)


# ----------------------------------------------------------------------------------------- sampling
_S = InfectionState


class MultiIndex_PopulationsArray(tuple):
    """Index (AgeGroup, InfectionState) of one compartment (the binding's MultiIndex of the populations array),
    accepted wherever populations take (group, state)."""

    def __new__(cls, group, state):
        return super().__new__(cls, (group, state))


def _copy_uv(dst, src):
    """UncertainValue assignment: value and (a copy of the) distribution."""
    dst.value = src.value
    d = src.get_distribution()
    dst._dist_set(copy.deepcopy(d) if d is not None else None)


def draw_sample_infection(model):
    """osecir::draw_sample_infection (models/ode_secir/parameter_space.h:153-195): seasonality and the
    parameters shared by all age groups are drawn once (group 0) and copied to every group; the age-specific
    ones are drawn per group."""
    p = model.parameters
    p.Seasonality.draw_sample()
    shared = ("TimeExposed", "TimeInfectedNoSymptoms", "RelativeTransmissionNoSymptoms",
              "RiskOfInfectionFromSymptomatic", "MaxRiskOfInfectionFromSymptomatic")
    for name in ("TimeExposed", "TimeInfectedNoSymptoms", "TimeInfectedSymptoms",
                 "RelativeTransmissionNoSymptoms", "RiskOfInfectionFromSymptomatic",
                 "MaxRiskOfInfectionFromSymptomatic"):
        getattr(p, name)[0].draw_sample()
    for i in range(p.get_num_groups()):
        for name in shared:
            arr = getattr(p, name)
            _copy_uv(arr[i], arr[0])
        for name in ("TimeInfectedSymptoms", "TimeInfectedSevere", "TimeInfectedCritical",
                     "TransmissionProbabilityOnContact", "RecoveredPerInfectedNoSymptoms", "DeathsPerCritical",
                     "SeverePerInfectedSymptoms", "CriticalPerSevere"):
            getattr(p, name)[i].draw_sample()


def draw_sample_demographics(model):
    """osecir::draw_sample_demographics (parameter_space.h:120-151): the capacities and the initial
    compartments E, INS, ISy, ISev, ICr, R are drawn; S absorbs the difference so each group keeps its total."""
    p, pop = model.parameters, model.populations
    p.ICUCapacity.draw_sample()
    p.TestAndTraceCapacity.draw_sample()
    for i in range(p.get_num_groups()):
        total = pop.get_group_total_AgeGroup(i)
        for st in (_S.Exposed, _S.InfectedNoSymptoms, _S.InfectedSymptoms, _S.InfectedSevere,
                   _S.InfectedCritical, _S.Recovered):
            pop[i, st].draw_sample()
        pop.set_difference_from_group_total_AgeGroup((i, _S.Susceptible), total)
        pop.set_difference_from_group_total_AgeGroup((i, _S.Susceptible), pop.get_group_total_AgeGroup(i))


def draw_sample(model_or_graph):
    """osecir::draw_sample (parameter_space.h:197-251). For a Model: infection and demographic parameters,
    then the contact matrices from the sampled dampings (`ContactPatterns.draw_sample`), then
    apply_constraints. For a ModelGraph: returns the sampled graph: node 0's infection parameters, contact
    dampings and dynamic NPIs are drawn once and given to every node (each keeps its own demographics, ICU and
    test-and-trace capacities and school holidays; its contact matrices are rebuilt from the sampled
    dampings, replacing the dampings they had); every edge gets the sampled contact dampings as mobility
    dampings and the node-0 dynamic NPIs. The graph passed in is changed too (sampled in place)."""
    if isinstance(model_or_graph, Model):
        m = model_or_graph
        draw_sample_infection(m)
        draw_sample_demographics(m)
        m.parameters.ContactPatterns.draw_sample()
        m.apply_constraints()
        return None
    graph = model_or_graph
    shared = graph.get_node(0).property
    draw_sample_infection(shared)
    shared_contacts = shared.parameters.ContactPatterns
    shared_contacts.draw_sample_dampings()
    shared_npis = shared.parameters.DynamicNPIsInfectedSymptoms
    shared_npis.draw_sample()
    sampled = ModelGraph()
    for v in range(graph.num_nodes):
        node = graph.get_node(v)
        m = node.property
        draw_sample_demographics(m)
        icu, tnt = UncertainValue(), UncertainValue()
        _copy_uv(icu, m.parameters.ICUCapacity)
        _copy_uv(tnt, m.parameters.TestAndTraceCapacity)
        holidays = list(m.parameters.ContactPatterns.school_holidays)
        m.parameters = copy.deepcopy(shared.parameters)
        _copy_uv(m.parameters.ICUCapacity, icu)
        _copy_uv(m.parameters.TestAndTraceCapacity, tnt)
        m.parameters.ContactPatterns.school_holidays = holidays
        m.parameters.ContactPatterns.make_matrix()
        m.apply_constraints()
        sampled.add_node(node.id, m)
    for e in graph._edges:
        ep = copy.deepcopy(e.property)
        n = ep.coefficients[0].baseline.size
        apply_dampings(ep.coefficients, shared_contacts.dampings, lambda g: make_mobility_damping_vector(n, g))
        ep.dynamic_npis_infected = copy.deepcopy(shared_npis)
        sampled.add_edge(e.start_node_idx, e.end_node_idx, ep)
    return sampled


def set_params_distributions_normal(model, t0, tmax, dev_rel):
    """osecir::set_params_distributions_normal (parameter_space.h:46-118): a truncated normal distribution
    around every sampled value within (1 -+ 2.6 dev_rel) * value, and a contact damping of value 0.5 at the
    middle of [t0, tmax] on every contact matrix. The reference passes
    dev_rel * value in the place of the quantile of ParameterDistributionNormal(lb, ub, mean, quantile) (its
    four-argument form); that is kept, so the distributions are the ones the reference samples."""
    big = np.finfo(np.float64).max
    dev_rel = float(dev_rel)

    def set_distribution(uv, min_val=0.001):
        v = uv.value
        lb = min(max(min_val, (1 - dev_rel * 2.6) * v), 0.1 * big)
        ub = min(max(min_val, (1 + dev_rel * 2.6) * v), 0.5 * big)
        diff = abs(lb - ub)
        if diff <= 1e-12 or diff <= max(abs(lb), abs(ub)) * np.finfo(np.float64).tiny:
            uv.set_distribution(ParameterDistributionConstant(min(max(min_val, v), 0.3 * big)))
        else:
            uv.set_distribution(ParameterDistributionNormal(lb, ub, min(max(min_val, v), 0.3 * big),
                                                            min(max(min_val, dev_rel * v), big)))

    p = model.parameters
    set_distribution(p.Seasonality, 0.0)
    set_distribution(p.ICUCapacity)
    set_distribution(p.TestAndTraceCapacity)
    for i in range(p.get_num_groups()):
        for st in _S:
            if st in (_S.Susceptible, _S.Dead, _S.Count):
                continue
            set_distribution(model.populations[i, st], 0.0)
    for i in range(p.get_num_groups()):
        for name in ("TimeExposed", "TimeInfectedNoSymptoms", "TimeInfectedSymptoms", "TimeInfectedSevere",
                     "TimeInfectedCritical", "TransmissionProbabilityOnContact", "RelativeTransmissionNoSymptoms",
                     "RecoveredPerInfectedNoSymptoms", "RiskOfInfectionFromSymptomatic",
                     "MaxRiskOfInfectionFromSymptomatic", "DeathsPerCritical", "SeverePerInfectedSymptoms",
                     "CriticalPerSevere"):
            set_distribution(getattr(p, name)[i])
    cp = p.ContactPatterns
    cp.dampings.append(DampingSampling(0.5, 0, 0, float(t0) + (float(tmax) - float(t0)) / 2,
                                       list(range(cp.cont_freq_mat.num_matrices)), np.ones(p.get_num_groups())))
    # the reference gives the distribution to the first contact damping sampling (the new one only if there
    # was none before)
    set_distribution(cp.dampings[0].value, 0.0)


class ParameterStudy:
    """ParameterStudy of one model (bindings compartments/parameter_studies.h:55-88): every run draws a
    sample of a copy of the model (draw_sample) and simulates it from t0 to tmax with step dt."""

    def __init__(self, model, t0, tmax, dt, num_runs):
        self.parameters = model
        self.t0, self.tmax, self.dt = float(t0), float(tmax), float(dt)
        self.num_runs = int(num_runs)

    def run(self, handle_result_func=None, integrator=None):
        sims = []
        for run in range(self.num_runs):
            m = copy.deepcopy(self.parameters)
            draw_sample(m)
            sim = Simulation(m, self.t0, self.dt, integrator)
            sim.advance(self.tmax)
            if handle_result_func is not None:
                handle_result_func(sim, run)
            else:
                sims.append(sim)
        return None if handle_result_func is not None else sims


class GraphParameterStudy:
    """ParameterStudy over a ModelGraph (parameter_studies.h:156-217, bindings :92-124): every run draws a
    sample of a copy of the graph (draw_sample: node 0's infection parameters, contact dampings and dynamic
    NPIs for all nodes, per-node demographics) and builds the simulation graph with node step = graph step =
    dt (make_sampled_graph_simulation). An optional sample(graph, run_idx) callback then edits the sampled
    graph; all runs execute as one device batch (mgo_ensemble_run)."""

    def __init__(self, graph, t0, tmax, dt, num_runs):
        self.parameters = graph
        self.t0, self.tmax, self.dt = float(t0), float(tmax), float(dt)
        self.num_runs = int(num_runs)

    def run(self, handle_result_func=None, sample=None, integrator=None):
        graphs, specs, cfgs = [], [], []
        for run in range(self.num_runs):
            g = draw_sample(copy.deepcopy(self.parameters))
            if sample is not None:
                sample(g, run)
            sg = MobilityGraph()
            for v in range(g.num_nodes):
                nd = g.get_node(v)
                sg.add_node(nd.id, nd.property, self.t0, self.dt)
            for e in g._edges:
                sg.add_edge(e.start_node_idx, e.end_node_idx, e.property)
            edges = [(e.start_node_idx, e.end_node_idx, e.property) for e in sg._edges]
            specs.append(_spec_from([sg.get_node(v).property.model for v in range(sg.num_nodes)], edges))
            cfgs.append(capi.sim_cfg(t0=self.t0, dt_graph=self.dt, dt_node=self.dt,
                                     integrator=(integrator or ControlledStepperWrapper()).cfg()))
            graphs.append(sg)
        outs, _ = _context().ensemble_run(specs, cfgs, [self.tmax] * self.num_runs)
        d0 = int(np.ceil(self.t0 - 1e-10))
        for sg, daily in zip(graphs, outs):
            _fill_results(sg, daily, d0)
        if handle_result_func is None:
            return graphs
        for run, sg in enumerate(graphs):
            handle_result_func(sg, run)
        return None


# ----------------------------------------------------------------------------------------- ensemble results
def interpolate_simulation_result(result, abs_tol=1e-10):
    """interpolate_simulation_result of a TimeSeries, or of a MobilityGraph: the node results at the integer
    days, one TimeSeries per node (data/analyze_result.h:186-265; binding osecir.cpp:224-227)."""
    if isinstance(result, _Graph):
        return [_interp_ts(result.get_node(v).property.result, abs_tol) for v in range(result.num_nodes)]
    return _interp_ts(result, abs_tol)


def interpolate_ensemble_results(ensemble):
    """interpolate_ensemble_results (analyze_result.h:77-86; binding osecir.cpp:228): [run][node] TimeSeries
    of the MobilityGraphs of a study."""
    return [interpolate_simulation_result(g) for g in ensemble]


def _ensemble_values(ensemble):
    """[run][node] TimeSeries as values[run][node][time][element]; the input must be uniform as
    interpolate_ensemble_results returns it (same nodes, time points and elements in every run)."""
    runs = len(ensemble)
    if runs == 0 or len(ensemble[0]) == 0:
        raise ValueError("ensemble results are empty")
    nodes = len(ensemble[0])
    T = ensemble[0][0].get_num_time_points()
    E = ensemble[0][0].get_num_elements()
    if T == 0 or E == 0:
        raise ValueError("ensemble results have no values")
    v = np.empty((runs, nodes, T, E))
    for r, run in enumerate(ensemble):
        if len(run) != nodes:
            raise ValueError("ensemble results not uniform: runs differ in their number of nodes")
        for n, ts in enumerate(run):
            if ts.get_num_time_points() != T or ts.get_num_elements() != E:
                raise ValueError("ensemble results not uniform: time points or elements differ")
            v[r, n] = np.stack(ts.values)
    return v


def _series(times, values):
    ts = TimeSeries(values.shape[-1])
    for t, row in zip(times, values):
        ts.add_time_point(t, row)
    return ts


def ensemble_mean(ensemble):
    """ensemble_mean (analyze_result.h:289-315; binding osecir.cpp:93): per node the mean over runs, summed in
    run order as value / num_runs, computed on the device (mgo_ensemble_mean_values); the times are the last
    run's, as in the reference."""
    v = _ensemble_values(ensemble)
    m = _context().ensemble_mean_values(v)
    return [_series(ensemble[-1][n].times, m[n]) for n in range(v.shape[1])]


def ensemble_percentile(ensemble, p):
    """ensemble_percentile (analyze_result.h:316-343; binding osecir.cpp:94): per node, time and element the
    sorted run values at index floor(num_runs * p), p in (0, 1), computed on the device
    (mgo_ensemble_percentiles_values); the times are the first run's."""
    v = _ensemble_values(ensemble)
    q = _context().ensemble_percentiles_values(v, [float(p)])[0]
    return [_series(ensemble[0][n].times, q[n]) for n in range(v.shape[1])]


# ----------------------------------------------------------------------------------------- graph files
def write_graph(graph, directory):
    """osecir.write_graph (binding io/mobility_io.h:34-45 over mio::write_graph, mobility_io.h:70-122): the
    ModelGraph as JSON files GraphNode<i>.json / GraphEdges_node<i>.json in `directory`."""
    from ._json import IOF_None
    from ._json import write_graph as _write
    _write(graph, directory, IOF_None)


def read_graph(directory):
    """osecir.read_graph (binding io/mobility_io.h:47-58 over mio::read_graph, mobility_io.h:124-192): the
    ModelGraph written by write_graph (or by the reference). A missing directory or file raises ValueError,
    a malformed file RuntimeError."""
    from ._json import read_graph as _read
    return _read(directory, 0, True)
