"""Host-side graph description: numpy arrays in the layout of include/mgo.h's mgo_graph_desc.

A GraphSpec is the structure-of-arrays form of a
``Graph<SimulationNode<osecir::Simulation>, MobilityEdge>`` (cpp/memilio/mobility/graph.h:144-303):
per-node osecir parameters (cpp/models/ode_secir/parameters.h:46-394), one contact matrix per node
with its dampings (cpp/memilio/epidemiology/contact_matrix.h, damping.h), initial populations and
the mobility edges with their coefficient vectors.
"""
import ctypes as C

import numpy as np

from . import capi

NC = capi.NUM_COMPARTMENTS

# osecir::InfectionState order (cpp/models/ode_secir/infection_state.h:33-47)
S, E, INS, INSC, ISY, ISYC, ISEV, ICR, R, D = range(NC)
COMPARTMENT_NAMES = ["S", "E", "INS", "INSC", "ISy", "ISyC", "ISev", "ICr", "R", "D"]


def default_scalars():
    """Reference defaults of the scalar parameters (parameters.h:46-394, 764-767)."""
    s = np.zeros(capi.NUM_SCALARS)
    s[capi.P_START_DAY] = 0.0
    s[capi.P_SEASONALITY] = 0.0
    s[capi.P_ICU_CAPACITY] = capi.DBL_MAX
    s[capi.P_TEST_AND_TRACE_CAPACITY] = capi.DBL_MAX
    s[capi.P_TEST_AND_TRACE_MAX_RISK] = 5.0
    return s


def default_age_params(G):
    """Reference defaults of the per-age parameters (parameters.h:96-324)."""
    a = np.zeros((capi.NUM_AGE_PARAMS, G))
    for k in (capi.A_TIME_EXPOSED, capi.A_TIME_INFECTED_NO_SYMPTOMS, capi.A_TIME_INFECTED_SYMPTOMS,
              capi.A_TIME_INFECTED_SEVERE, capi.A_TIME_INFECTED_CRITICAL,
              capi.A_TRANSMISSION_PROBABILITY_ON_CONTACT, capi.A_RELATIVE_TRANSMISSION_NO_SYMPTOMS,
              capi.A_RISK_OF_INFECTION_FROM_SYMPTOMATIC, capi.A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC):
        a[k, :] = 1.0
    return a


class GraphSpec:
    """Arrays of one graph; ``desc()`` returns a capi.GraphDesc pointing into them."""

    def __init__(self, num_nodes, G):
        self.num_nodes = int(num_nodes)
        self.G = int(G)
        n = self.num_nodes
        self.node_scalars = np.tile(default_scalars(), (n, 1))
        self.age_params = np.tile(default_age_params(G)[None], (n, 1, 1))
        self.contact_baseline = np.zeros((n, G, G))
        self.contact_minimum = None
        self.damping_offsets = None
        self.damping_times = None
        self.damping_levels = None
        self.damping_types = None
        self.damping_coeffs = None
        self.populations = np.zeros((n, G, NC))
        self.edge_start = np.zeros(0, np.int32)
        self.edge_end = np.zeros(0, np.int32)
        self.edge_coeffs = np.zeros(0)   # [E] (one value per edge) or [E][10G]
        self.npis = None                 # dynamic NPI arrays (set_dynamic_npis)
        self.edge_coeff_minimum = None   # [E] or [E][10G] minimum mobility coefficients
        self.out_dampings = None         # per-node out-edge damping arrays (set_out_dampings)
        self.out_npis = None             # per-node out-edge NPI arrays (set_out_dynamic_npis)
        self._desc = None

    # ---------------------------------------------------------------- building
    def set_dampings(self, per_node):
        """per_node: list (len num_nodes) of lists of (time, level, type, coeffs[G][G] or scalar)."""
        G = self.G
        offs = [0]
        times, levels, types, coeffs = [], [], [], []
        for lst in per_node:
            for (t, lev, typ, c) in lst:
                times.append(float(t))
                levels.append(int(lev))
                types.append(int(typ))
                coeffs.append(np.broadcast_to(np.asarray(c, dtype=np.float64), (G, G)))
            offs.append(len(times))
        self.damping_offsets = np.asarray(offs, np.int32)
        self.damping_times = np.asarray(times, np.float64)
        self.damping_levels = np.asarray(levels, np.int32)
        self.damping_types = np.asarray(types, np.int32)
        self.damping_coeffs = (np.stack(coeffs) if coeffs else np.zeros((0, G, G))).astype(np.float64)
        self._desc = None

    def set_dynamic_npis(self, per_node):
        """Dynamic NPIs of the nodes' contact matrices (osecir DynamicNPIsInfectedSymptoms).

        per_node: list (len num_nodes) of None or dicts with keys ``thresholds`` (list of (value, dampings),
        dampings = list of (value, level, type, group_weights[G])), and optionally ``duration`` (14),
        ``delay`` (0), ``base_value`` (1), ``directive_begin`` (lowest), ``directive_end`` (max)."""
        self.npis = self._npi_arrays(per_node)
        self._desc = None

    def set_out_dynamic_npis(self, per_node):
        """Dynamic NPIs of every out-edge of a node (MobilityParameters::set_dynamic_npis_infected); same
        per-node form as set_dynamic_npis (apply_mobility uses duration and base_value)."""
        self.out_npis = self._npi_arrays(per_node)
        self._desc = None

    def set_out_dampings(self, per_node):
        """Dampings of the mobility coefficients of every out-edge of a node: per node a list of
        (time, level, type, values[G]) (the value of group g applies to its 10 compartments)."""
        G = self.G
        offs, times, levels, types, coeffs = [0], [], [], [], []
        for lst in per_node:
            for (t, lev, typ, c) in (lst or []):
                times.append(float(t))
                levels.append(int(lev))
                types.append(int(typ))
                coeffs.append(np.broadcast_to(np.asarray(c, dtype=np.float64), (G,)))
            offs.append(len(times))
        self.out_dampings = {"offsets": np.asarray(offs, np.int32), "times": np.asarray(times, np.float64),
                             "levels": np.asarray(levels, np.int32), "types": np.asarray(types, np.int32),
                             "coeffs": np.ascontiguousarray(np.stack(coeffs) if coeffs else np.zeros((0, G)))}
        self._desc = None

    def _npi_arrays(self, per_node):
        G = self.G
        node_offsets, params, thresholds, damp_offsets = [0], [], [], [0]
        values, levels, types, weights = [], [], [], []
        for cfg in per_node:
            cfg = cfg or {}
            p = np.zeros(capi.NPI_NUM_PARAMS)
            p[capi.NPI_DURATION] = cfg.get("duration", 14.0)
            p[capi.NPI_DELAY] = cfg.get("delay", 0.0)
            p[capi.NPI_BASE_VALUE] = cfg.get("base_value", 1.0)
            p[capi.NPI_DIRECTIVE_BEGIN] = cfg.get("directive_begin", -capi.DBL_MAX)
            p[capi.NPI_DIRECTIVE_END] = cfg.get("directive_end", capi.DBL_MAX)
            params.append(p)
            for value, dampings in cfg.get("thresholds", []):
                thresholds.append(float(value))
                for (v, lev, typ, w) in dampings:
                    values.append(float(v))
                    levels.append(int(lev))
                    types.append(int(typ))
                    weights.append(np.broadcast_to(np.asarray(w, dtype=np.float64), (G,)))
                damp_offsets.append(len(values))
            node_offsets.append(len(thresholds))
        return {
            "node_offsets": np.asarray(node_offsets, np.int32),
            "node_params": np.ascontiguousarray(np.stack(params), dtype=np.float64),
            "thresholds": np.asarray(thresholds, np.float64),
            "damping_offsets": np.asarray(damp_offsets, np.int32),
            "damping_values": np.asarray(values, np.float64),
            "damping_levels": np.asarray(levels, np.int32),
            "damping_types": np.asarray(types, np.int32),
            "damping_weights": np.ascontiguousarray(np.stack(weights) if weights else np.zeros((0, G))),
        }

    def set_edges(self, start, end, coeffs):
        """coeffs: [E] (one coefficient for every compartment of the edge) or [E][10G]."""
        self.edge_start = np.ascontiguousarray(start, dtype=np.int32)
        self.edge_end = np.ascontiguousarray(end, dtype=np.int32)
        c = np.asarray(coeffs, dtype=np.float64)
        assert c.ndim == 1 or c.shape[1] == NC * self.G
        self.edge_coeffs = np.ascontiguousarray(c)
        self._desc = None

    def edge_coeff_matrix(self):
        """[E][10G] view of the coefficients."""
        c = self.edge_coeffs
        return np.repeat(c[:, None], NC * self.G, axis=1) if c.ndim == 1 else c

    @property
    def num_edges(self):
        return int(self.edge_start.shape[0])

    # ---------------------------------------------------------------- C view
    def _finalize(self):
        for name in ("node_scalars", "age_params", "contact_baseline", "contact_minimum", "damping_times",
                     "damping_coeffs", "populations", "edge_coeffs"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, np.ascontiguousarray(v, dtype=np.float64))
        for name in ("damping_offsets", "damping_levels", "damping_types", "edge_start", "edge_end"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, np.ascontiguousarray(v, dtype=np.int32))

    def desc(self):
        if self._desc is None:
            self._finalize()
            d = capi.GraphDesc()
            d.num_nodes = self.num_nodes
            d.num_groups = self.G
            d.node_scalars = capi.dptr(self.node_scalars)
            d.age_params = capi.dptr(self.age_params)
            d.contact_baseline = capi.dptr(self.contact_baseline)
            d.contact_minimum = capi.dptr(self.contact_minimum)
            d.damping_offsets = capi.iptr(self.damping_offsets)
            d.damping_times = capi.dptr(self.damping_times)
            d.damping_levels = capi.iptr(self.damping_levels)
            d.damping_types = capi.iptr(self.damping_types)
            d.damping_coeffs = capi.dptr(self.damping_coeffs)
            d.populations = capi.dptr(self.populations)
            d.num_edges = self.num_edges
            d.edge_start = capi.iptr(self.edge_start) if self.num_edges else None
            d.edge_end = capi.iptr(self.edge_end) if self.num_edges else None
            d.edge_coeffs = capi.dptr(self.edge_coeffs) if self.num_edges else None
            d.edge_coeff_len = 1 if self.edge_coeffs.ndim == 1 else NC * self.G
            def npi_struct(n):
                return capi.DynamicNpis(
                    capi.iptr(n["node_offsets"]), capi.dptr(n["node_params"]), capi.dptr(n["thresholds"]),
                    capi.iptr(n["damping_offsets"]), capi.dptr(n["damping_values"]), capi.iptr(n["damping_levels"]),
                    capi.iptr(n["damping_types"]), capi.dptr(n["damping_weights"]))
            if self.npis is not None:
                self._npi_struct = npi_struct(self.npis)
                d.dynamic_npis = C.pointer(self._npi_struct)
            if self.out_npis is not None:
                self._out_npi_struct = npi_struct(self.out_npis)
                d.out_dynamic_npis = C.pointer(self._out_npi_struct)
            if self.edge_coeff_minimum is not None and self.num_edges:
                self.edge_coeff_minimum = np.ascontiguousarray(self.edge_coeff_minimum, dtype=np.float64)
                assert self.edge_coeff_minimum.shape == self.edge_coeffs.shape
                d.edge_coeff_minimum = capi.dptr(self.edge_coeff_minimum)
            if self.out_dampings is not None:
                o = self.out_dampings
                d.out_damping_offsets = capi.iptr(o["offsets"])
                d.out_damping_times = capi.dptr(o["times"])
                d.out_damping_levels = capi.iptr(o["levels"])
                d.out_damping_types = capi.iptr(o["types"])
                d.out_damping_coeffs = capi.dptr(o["coeffs"])
            self._desc = d
        return self._desc

    def state_size(self):
        return self.G * NC

    def num_days(self, t0, tmax):
        """Days of interpolate_simulation_result: ceil(t0 - 1e-10) .. floor(tmax + 1e-10)."""
        return int(np.floor(tmax + 1e-10)) - int(np.ceil(t0 - 1e-10)) + 1
