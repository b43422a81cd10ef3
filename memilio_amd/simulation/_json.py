"""Parameter-graph JSON files of the osecir Graph-ODE path: write_graph / read_graph.

Restates the reference's file layout (cpp/memilio/io/mobility_io.h:70-210): a directory with one
``GraphNode<i>.json`` per node ({"NodeId": id, "Model": model}) and one ``GraphEdges_node<i>.json`` per node
with out-edges ([{"StartNodeIndex", "EndNodeIndex", "Parameters"}]). The objects inside follow the
reference's serializers, JSON flavour (cpp/memilio/io/json_serializer.h, io.h:582-653):

  UncertainValue        {"Value": v[, "Distribution": {...}]}        utils/uncertain_value.h:195-204
  CustomIndexArray      {"Dimensions": d, "Elements": [...]}          utils/custom_index_array.h:745-750
                        d = n for one category, {"Indices": {"Element0": n0, "Element1": n1}} for two
                        (utils/index.h:166-169, 259-263; tuples as "Element<k>", io.h:582-595)
  Eigen matrix/vector   {"Rows": r, "Columns": c, "Elements": row-major}   io.h:620-653
  Damping               {"Time", "Type", "Level", "Coeffs"}           epidemiology/damping.h:238-245
  ContactMatrix         {"Baseline", "Minimum", "Dampings": [...]}    epidemiology/contact_matrix.h:211-217
  ContactMatrixGroup    {"Matrices": [...]}                           epidemiology/contact_matrix.h:438-442
  UncertainContactMatrix {"ContactMatrix"[, "SchoolHolidayDamping", "SchoolHolidays", "Dampings"]}
                                                                      epidemiology/uncertain_matrix.h:209-218
  DampingSampling       {"Time", "Type", "Level", "Value", "MatrixIndices", "GroupWeights"}
                                                                      epidemiology/damping_sampling.h:206-215
  DynamicNPIs           {"Thresholds": [{"Element0": t, "Element1": [...]}], "Duration", "Delay",
                         "BaseValue", "DirectiveBegin", "DirectiveEnd"}   epidemiology/dynamic_npis.h:206-216
  osecir Model          {"Parameters": {<tag name>: value, ...}, "Populations": array}
                                                                      models/ode_secir/model.h:219-225
  MobilityParameters    {"Coefficients": group, "DynamicNPIs": npis}  metapopulation_mobility_instant.h:259-264

Distributions of uncertain values become ParameterDistribution objects (Normal, Uniform, LogNormal,
Exponential, Constant; parameter_distributions.h:925-958). Reading is as strict as the reference's: a
missing element is an error.
"""
import json
import os

import numpy as np

IOF_None = 0
IOF_OmitDistributions = 1 << 0
IOF_OmitValues = 1 << 1
IOF_IncludeTypeInfo = 1 << 2


class JsonFormatError(RuntimeError):
    """A file that does not hold the expected object (the reference's StatusCode::KeyNotFound /
    InvalidType / OutOfRange)."""


# ------------------------------------------------------------------------------------------ elements
def _expect(js, key, where):
    if not isinstance(js, dict) or key not in js:
        raise JsonFormatError(f"{where}: element {key!r} missing")
    return js[key]


def _number(v, where):
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise JsonFormatError(f"{where}: number expected")
    return float(v)


def _integer(v, where):
    if isinstance(v, bool) or not isinstance(v, int):
        raise JsonFormatError(f"{where}: integer expected")
    return int(v)


def uv_to_json(value, flags=IOF_None, distribution=None):
    from . import distribution_to_json
    js = {}
    if not flags & IOF_OmitValues:
        js["Value"] = float(value)
    if not flags & IOF_OmitDistributions and distribution is not None:
        js["Distribution"] = distribution_to_json(distribution)
    return js


def uv_from_json(js, flags=IOF_None, where="UncertainValue"):
    """(value, distribution or None) (uncertain_value.h:211-250)."""
    if not isinstance(js, dict):
        raise JsonFormatError(f"{where}: object expected")
    from . import distribution_from_json
    v = 0.0 if flags & IOF_OmitValues else _number(_expect(js, "Value", where), where)
    dist = None
    if not flags & (IOF_OmitValues | IOF_OmitDistributions) and js.get("Distribution") is not None:
        try:
            dist = distribution_from_json(js["Distribution"])
        except (KeyError, ValueError, TypeError) as e:
            raise JsonFormatError(f"{where}: distribution: {e}") from None
    return v, dist


def matrix_to_json(a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    return {"Rows": int(a.shape[0]), "Columns": int(a.shape[1]), "Elements": [float(x) for x in a.ravel()]}


def matrix_from_json(js, where="Matrix"):
    r = _integer(_expect(js, "Rows", where), where)
    c = _integer(_expect(js, "Columns", where), where)
    el = _expect(js, "Elements", where)
    if not isinstance(el, list) or len(el) != r * c:
        raise JsonFormatError(f"{where}: {r}x{c} matrix needs {r * c} elements")
    return np.array([_number(x, where) for x in el], dtype=np.float64).reshape(r, c)


def vector_from_json(js, where="Vector"):
    m = matrix_from_json(js, where)
    if m.shape[1] != 1:
        raise JsonFormatError(f"{where}: column vector expected")
    return m[:, 0].copy()


def array1_to_json(values, flags=IOF_None, dists=None):
    """CustomIndexArray<UncertainValue, AgeGroup>."""
    dists = dists or {}
    return {"Dimensions": len(values),
            "Elements": [uv_to_json(v, flags, dists.get(i)) for i, v in enumerate(values)]}


def array1_from_json(js, flags=IOF_None, where="Array"):
    n = _integer(_expect(js, "Dimensions", where), where)
    el = _expect(js, "Elements", where)
    if not isinstance(el, list) or len(el) != n:
        raise JsonFormatError(f"{where}: Dimensions of Array don't match the number of elements.")
    vals, dists = np.zeros(n), {}
    for i, e in enumerate(el):
        vals[i], d = uv_from_json(e, flags, where)
        if d is not None:
            dists[i] = d
    return vals, dists


def array2_to_json(values, flags=IOF_None, dists=None):
    """CustomIndexArray<UncertainValue, AgeGroup, InfectionState> (Populations), flat index g * 10 + c."""
    dists = dists or {}
    n0, n1 = values.shape
    return {"Dimensions": {"Indices": {"Element0": int(n0), "Element1": int(n1)}},
            "Elements": [uv_to_json(v, flags, dists.get(i)) for i, v in enumerate(values.ravel())]}


def array2_from_json(js, flags=IOF_None, where="Populations"):
    dims = _expect(js, "Dimensions", where)
    idx = _expect(dims, "Indices", where)
    n0 = _integer(_expect(idx, "Element0", where), where)
    n1 = _integer(_expect(idx, "Element1", where), where)
    el = _expect(js, "Elements", where)
    if not isinstance(el, list) or len(el) != n0 * n1:
        raise JsonFormatError(f"{where}: Dimensions of Array don't match the number of elements.")
    vals, dists = np.zeros(n0 * n1), {}
    for i, e in enumerate(el):
        vals[i], d = uv_from_json(e, flags, where)
        if d is not None:
            dists[i] = d
    return vals.reshape(n0, n1), dists


def _pair(a, b):
    return {"Element0": a, "Element1": b}


# ------------------------------------------------------------------------------------------ dampings
def damping_to_json(time, level, type, coeffs):
    return {"Time": float(time), "Type": int(type), "Level": int(level), "Coeffs": matrix_to_json(coeffs)}


def damping_from_json(js, where="Damping"):
    return (_number(_expect(js, "Time", where), where), _integer(_expect(js, "Level", where), where),
            _integer(_expect(js, "Type", where), where), matrix_from_json(_expect(js, "Coeffs", where), where))


def damping_sampling_to_json(ds, flags=IOF_None):
    return {"Time": float(ds.time), "Type": int(ds.type), "Level": int(ds.level),
            "Value": uv_to_json(ds.value.value, flags, ds.value.get_distribution()),
            "MatrixIndices": [int(k) for k in ds.matrix_indices],
            "GroupWeights": matrix_to_json(np.asarray(ds.group_weights, dtype=np.float64).ravel())}


def damping_sampling_from_json(js, flags=IOF_None, where="DampingSampling"):
    from . import DampingSampling
    v, dist = uv_from_json(_expect(js, "Value", where), flags, where)
    idx = _expect(js, "MatrixIndices", where)
    if not isinstance(idx, list):
        raise JsonFormatError(f"{where}: MatrixIndices must be a list")
    ds = DampingSampling(v, _integer(_expect(js, "Level", where), where), _integer(_expect(js, "Type", where), where),
                         _number(_expect(js, "Time", where), where), [_integer(k, where) for k in idx],
                         vector_from_json(_expect(js, "GroupWeights", where), where))
    if dist is not None:
        ds.value._dist_set(dist)
    return ds


def dynamic_npis_to_json(dn, flags=IOF_None):
    return {"Thresholds": [_pair(float(t), [damping_sampling_to_json(d, flags) for d in ds])
                           for t, ds in dn.threshold],
            "Duration": float(dn.duration), "Delay": float(dn.implementation_delay),
            "BaseValue": float(dn.base_value), "DirectiveBegin": float(dn.directive_begin),
            "DirectiveEnd": float(dn.directive_end)}


def dynamic_npis_from_json(js, flags=IOF_None, where="DynamicNPIs"):
    """DynamicNPIs::deserialize (dynamic_npis.h:222-242): thresholds inserted with set_threshold."""
    from . import DynamicNPIs
    dn = DynamicNPIs()
    ths = _expect(js, "Thresholds", where)
    if not isinstance(ths, list):
        raise JsonFormatError(f"{where}: Thresholds must be a list")
    dn.duration = _number(_expect(js, "Duration", where), where)
    dn.implementation_delay = _number(_expect(js, "Delay", where), where)
    dn.base_value = _number(_expect(js, "BaseValue", where), where)
    for e in ths:
        lst = _expect(e, "Element1", where)
        if not isinstance(lst, list):
            raise JsonFormatError(f"{where}: threshold dampings must be a list")
        dn.set_threshold(_number(_expect(e, "Element0", where), where),
                         [damping_sampling_from_json(d, flags, where) for d in lst])
    dn.directive_begin = _number(_expect(js, "DirectiveBegin", where), where)
    dn.directive_end = _number(_expect(js, "DirectiveEnd", where), where)
    return dn


# ------------------------------------------------------------------------------------------ contact matrices
def contact_matrix_to_json(cm):
    G = cm.num_groups
    return {"Baseline": matrix_to_json(cm.baseline), "Minimum": matrix_to_json(cm.minimum),
            "Dampings": [damping_to_json(d.time, d.level, d.type, np.broadcast_to(d.coeffs, (G, G)))
                         for d in cm.dampings]}


def contact_matrix_from_json(js, where="ContactMatrix"):
    from . import ContactMatrix, Damping
    b = matrix_from_json(_expect(js, "Baseline", where), where)
    m = matrix_from_json(_expect(js, "Minimum", where), where)
    cm = ContactMatrix(b, m)
    lst = _expect(js, "Dampings", where)
    if not isinstance(lst, list):
        raise JsonFormatError(f"{where}: Dampings must be a list")
    for d in lst:
        t, lev, typ, c = damping_from_json(d, where)
        cm.add_damping(Damping(c, t, lev, typ))
    return cm


def uncertain_contact_matrix_to_json(ucm, flags=IOF_None):
    js = {"ContactMatrix": {"Matrices": [contact_matrix_to_json(cm) for cm in ucm.cont_freq_mat.matrices]}}
    if not flags & IOF_OmitDistributions:
        js["SchoolHolidayDamping"] = damping_sampling_to_json(ucm.school_holiday_damping, flags)
        js["SchoolHolidays"] = [_pair(float(a), float(b)) for a, b in ucm.school_holidays]
        js["Dampings"] = [damping_sampling_to_json(d, flags) for d in ucm.dampings]
    return js


def uncertain_contact_matrix_from_json(js, flags=IOF_None, where="ContactPatterns"):
    from . import ContactMatrixGroup, UncertainContactMatrix
    mats = _expect(_expect(js, "ContactMatrix", where), "Matrices", where)
    if not isinstance(mats, list) or not mats:
        raise JsonFormatError(f"{where}: Matrices must be a non-empty list")
    cms = [contact_matrix_from_json(m, where) for m in mats]
    grp = ContactMatrixGroup(len(cms), cms[0].num_groups)
    for k, cm in enumerate(cms):
        grp[k] = cm
    ucm = UncertainContactMatrix(grp)
    if not flags & IOF_OmitDistributions:
        ucm.school_holiday_damping = damping_sampling_from_json(_expect(js, "SchoolHolidayDamping", where), flags,
                                                                where)
        hol = _expect(js, "SchoolHolidays", where)
        ds = _expect(js, "Dampings", where)
        if not isinstance(hol, list) or not isinstance(ds, list):
            raise JsonFormatError(f"{where}: SchoolHolidays and Dampings must be lists")
        ucm.school_holidays = [(_number(_expect(h, "Element0", where), where),
                                _number(_expect(h, "Element1", where), where)) for h in hol]
        ucm.dampings = [damping_sampling_from_json(d, flags, where) for d in ds]
    return ucm


# ------------------------------------------------------------------------------------------ model, edges
def model_to_json(model, flags=IOF_None):
    from .osecir import _AGES, _SCALARS
    p = model.parameters
    dists = p._dist
    par = {"StartDay": float(p._sc[_SCALARS["StartDay"]])}
    for name, k in _SCALARS.items():
        if name != "StartDay":
            par[name] = uv_to_json(p._sc[k], flags, dists.get(name))
    for name, k in _AGES.items():
        par[name] = array1_to_json(p._age[k], flags, dists.get(name))
    par["ContactPatterns"] = uncertain_contact_matrix_to_json(p.ContactPatterns, flags)
    par["DynamicNPIsInfectedSymptoms"] = dynamic_npis_to_json(p.DynamicNPIsInfectedSymptoms, flags)
    return {"Parameters": par, "Populations": array2_to_json(model.populations._v, flags, model.populations._dist)}


def model_from_json(js, flags=IOF_None, where="Model"):
    from .osecir import _AGES, _SCALARS, Model
    par = _expect(js, "Parameters", where)
    pops, pdist = array2_from_json(_expect(js, "Populations", where), flags, where + ".Populations")
    G, nc = pops.shape
    if nc != 10:
        raise JsonFormatError(f"{where}: osecir populations need 10 compartments per age group, got {nc}")
    m = Model(G)
    p = m.parameters
    m.populations._v[:] = pops
    m.populations._dist = pdist
    p._sc[_SCALARS["StartDay"]] = _number(_expect(par, "StartDay", where), where)
    for name, k in _SCALARS.items():
        if name != "StartDay":
            p._sc[k], d = uv_from_json(_expect(par, name, where), flags, f"{where}.{name}")
            if d is not None:
                p._dist[name] = d
    for name, k in _AGES.items():
        vals, d = array1_from_json(_expect(par, name, where), flags, f"{where}.{name}")
        if vals.size != G:
            raise JsonFormatError(f"{where}.{name}: {vals.size} age groups, populations have {G}")
        p._age[k][:] = vals
        if d:
            p._dist[name] = d
    ucm = uncertain_contact_matrix_from_json(_expect(par, "ContactPatterns", where), flags, where + ".ContactPatterns")
    if ucm.cont_freq_mat[0].num_groups != G:
        raise JsonFormatError(f"{where}.ContactPatterns: matrices of size {ucm.cont_freq_mat[0].num_groups}, "
                              f"populations have {G} age groups")
    p.ContactPatterns = ucm
    p.DynamicNPIsInfectedSymptoms = dynamic_npis_from_json(_expect(par, "DynamicNPIsInfectedSymptoms", where), flags,
                                                           where + ".DynamicNPIsInfectedSymptoms")
    return m


def mobility_parameters_to_json(mp, flags=IOF_None):
    mats = []
    for c in mp.coefficients.matrices:
        n = c.baseline.size
        mats.append({"Baseline": matrix_to_json(c.baseline), "Minimum": matrix_to_json(c.minimum),
                     "Dampings": [damping_to_json(d.time, d.level, d.type, np.broadcast_to(d.coeffs, (n,)))
                                  for d in c.dampings]})
    return {"Coefficients": {"Matrices": mats}, "DynamicNPIs": dynamic_npis_to_json(mp.dynamic_npis_infected, flags)}


def mobility_parameters_from_json(js, flags=IOF_None, where="MobilityParameters"):
    from . import MobilityCoefficientGroup, MobilityCoefficients, MobilityDamping, MobilityParameters
    mats = _expect(_expect(js, "Coefficients", where), "Matrices", where)
    if not isinstance(mats, list) or not mats:
        raise JsonFormatError(f"{where}: Matrices must be a non-empty list")
    grp = MobilityCoefficientGroup(len(mats), 1)
    for k, m in enumerate(mats):
        c = MobilityCoefficients(vector_from_json(_expect(m, "Baseline", where), where),
                                 vector_from_json(_expect(m, "Minimum", where), where))
        lst = _expect(m, "Dampings", where)
        if not isinstance(lst, list):
            raise JsonFormatError(f"{where}: Dampings must be a list")
        for d in lst:
            t, lev, typ, coeffs = damping_from_json(d, where)
            c.add_damping(MobilityDamping(coeffs.ravel(), t, lev, typ))
        grp[k] = c
    mp = MobilityParameters(grp)
    mp.dynamic_npis_infected = dynamic_npis_from_json(_expect(js, "DynamicNPIs", where), flags, where + ".DynamicNPIs")
    return mp


# ------------------------------------------------------------------------------------------ graph files
def write_graph(graph, directory, ioflags=IOF_None):
    """mio::write_graph (mobility_io.h:70-122): GraphNode<i>.json for every node, GraphEdges_node<i>.json for
    every node with out-edges; the directory is created if missing."""
    if graph.num_nodes == 0:
        raise ValueError("write_graph: the graph has no nodes")
    os.makedirs(directory, exist_ok=True)
    for i in range(graph.num_nodes):
        node = graph.get_node(i)
        with open(os.path.join(directory, f"GraphNode{i}.json"), "w") as f:
            json.dump({"NodeId": int(node.id), "Model": model_to_json(node.property, ioflags)}, f, indent=1)
        out = [e for e in graph._edges if e.start_node_idx == i]
        if out:
            with open(os.path.join(directory, f"GraphEdges_node{i}.json"), "w") as f:
                json.dump([{"StartNodeIndex": int(e.start_node_idx), "EndNodeIndex": int(e.end_node_idx),
                            "Parameters": mobility_parameters_to_json(e.property, ioflags)} for e in out], f,
                          indent=1)


def _read_json(path):
    try:
        with open(path) as f:
            return json.load(f)
    except FileNotFoundError:
        raise ValueError(f"{path}: no such file") from None
    except json.JSONDecodeError as e:
        raise JsonFormatError(f"{path}: {e}") from None


def read_graph(directory, ioflags=IOF_None, read_edges=True):
    """mio::read_graph (mobility_io.h:124-192): nodes GraphNode0.json, GraphNode1.json, ... until the first
    missing index; with read_edges, GraphEdges_node<i>.json of every node is read (a node without the file
    is an error, as in the reference, whose write_graph writes none for a node without out-edges). A missing
    directory or file raises ValueError, a malformed one JsonFormatError."""
    from .osecir import ModelGraph
    if not os.path.isdir(directory):
        raise ValueError(f"Directory {directory} does not exist.")
    graph = ModelGraph()
    i = 0
    while True:
        path = os.path.join(directory, f"GraphNode{i}.json")
        if not os.path.exists(path):
            break
        js = _read_json(path)
        if not isinstance(js, dict) or isinstance(js.get("NodeId"), bool) or not isinstance(js.get("NodeId"), int):
            raise JsonFormatError(f"{path}, NodeId must be an integer.")
        graph._nodes.append(_node(js["NodeId"], model_from_json(_expect(js, "Model", path), ioflags, path)))
        i += 1
    if read_edges:
        for i in range(graph.num_nodes):
            path = os.path.join(directory, f"GraphEdges_node{i}.json")
            js = _read_json(path)
            if not isinstance(js, list):
                raise JsonFormatError(f"{path}: list of edges expected")
            for e in js:
                end = e.get("EndNodeIndex") if isinstance(e, dict) else None
                if isinstance(end, bool) or not isinstance(end, int) or end < 0:
                    raise JsonFormatError(f"{path}, EndNodeIndex must be an integer.")
                if end >= graph.num_nodes:
                    raise JsonFormatError(f"{path}, EndNodeIndex not in range of number of graph nodes.")
                graph.add_edge(i, end, mobility_parameters_from_json(_expect(e, "Parameters", path), ioflags, path))
    return graph


def _node(id, model):
    from .osecir import _Node
    return _Node(id, model)
