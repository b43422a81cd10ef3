"""The Python mirror of memilio.simulation.osecir on the device: the reference's Python mobility example,
a GraphParameterStudy batch and osecir.simulate, each against the oracle on the same descriptor."""
import copy

import numpy as np
import pytest

import memilio_amd.simulation as mio
import memilio_amd.simulation.osecir as osecir
from memilio_amd import capi
from tests import pymirror_setups as ps

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    assert a.shape == b.shape, (a.shape, b.shape)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6)))


def _spec(graph):
    nodes = [graph.get_node(v).property for v in range(graph.num_nodes)]
    edges = [(e.start_node_idx, e.end_node_idx, e.property) for e in graph._edges]
    return osecir._spec_from([n.model for n in nodes], edges)


def _results(graph):
    """[day][node][10G] of the node results."""
    return np.stack([np.stack(graph.get_node(v).property.result.values) for v in range(graph.num_nodes)], axis=1)


def test_mobility_simulation_example(oracle):
    graph = ps.mobility_example_graph()
    sim = osecir.MobilitySimulation(graph, 0.0, dt=0.5)
    sim.advance(50)
    assert sim.t == 50
    r0 = osecir.interpolate_simulation_result(sim.graph.get_node(0).property.result)
    assert r0.as_ndarray().shape == (11, 51)
    cfg = capi.sim_cfg(t0=0.0, dt_graph=0.5, dt_node=0.1)
    ref, _ = oracle.graph_run(_spec(graph), cfg, 50.0)
    assert rel_err(_results(sim.graph), ref) < 1e-6


def test_graph_parameter_study(oracle):
    """Every run simulates its own sample of the graph (draw_sample over the normal distributions of
    set_params_distributions_normal, then the sample callback); each run's device result is checked against
    the oracle on the graph that run sampled."""
    graph = ps.study_graph(n=5, G=2)
    for v in range(graph.num_nodes):
        osecir.set_params_distributions_normal(graph.get_node(v).property, 1.0, 30.0, 0.2)
    study = osecir.GraphParameterStudy(graph, t0=1, tmax=30, dt=0.5, num_runs=6)
    sampled_graphs, seen = {}, []

    def sample(g, run_idx):
        ps.scale_contacts(g, run_idx)
        sampled_graphs[run_idx] = copy.deepcopy(g)

    def handle(g, run_idx):
        seen.append(run_idx)
        assert g.get_node(0).property.result.get_time(0) == 1
        assert g.get_node(0).property.result.get_last_time() == 30
        sampled = sampled_graphs[run_idx]
        sg = osecir.MobilityGraph()
        for v in range(sampled.num_nodes):
            sg.add_node(v, sampled.get_node(v).property, 1.0, 0.5)
        for e in sampled._edges:
            sg.add_edge(e.start_node_idx, e.end_node_idx, e.property)
        cfg = capi.sim_cfg(t0=1.0, dt_graph=0.5, dt_node=0.5)
        ref, _ = oracle.graph_run(_spec(sg), cfg, 30.0)
        assert rel_err(_results(g), ref) < 1e-6

    mio.seed_random_number_generator(17)
    study.run(handle, sample=sample)
    assert seen == list(range(6))
    # the runs are samples: their shared infection parameters and initial compartments differ
    t_exp = {sampled_graphs[r].get_node(0).property.parameters.TimeExposed[0].value for r in range(6)}
    e0 = {sampled_graphs[r].get_node(3).property.populations[0, osecir.InfectionState.Exposed].value
          for r in range(6)}
    assert len(t_exp) == 6 and len(e0) == 6


def test_simulate_single_model(oracle):
    model = ps.study_model(G=3, seed=4)
    core = osecir.RKIntegratorCore(1e-8, 1e-8, 1e-6, 1.0)
    ts = osecir.simulate(0.0, 40.0, 0.3, model, core)
    assert ts.get_num_time_points() == 41 and ts.get_last_time() == 40
    spec = osecir._spec_from([model], [])
    cfg = capi.sim_cfg(t0=0.0, dt_node=0.3, integrator=core.cfg())
    ref, _ = oracle.simulate_daily(spec, cfg, 40.0)
    assert rel_err(np.stack(ts.values), ref) < 1e-6


def test_mobility_simulation_dynamic_npis(oracle):
    """The reference's Python mobility example with DynamicNPIsInfectedSymptoms on every node."""
    graph = ps.mobility_example_graph()
    for v in range(graph.num_nodes):
        dn = graph.get_node(v).property.model.parameters.DynamicNPIsInfectedSymptoms
        dn.implementation_delay = 2.0
        dn.duration = 7.0
        dn.base_value = 100000
        dn.set_threshold(100.0, [mio.DampingSampling(0.5, 0, 0, 0.0, [0], np.ones(1))])
        dn.set_threshold(500.0, [mio.DampingSampling(0.3, 1, 0, 0.0, [0], np.ones(1))])
    spec = _spec(graph)
    sim = osecir.MobilitySimulation(graph, 0.0, dt=0.5)
    sim.advance(40)
    cfg = capi.sim_cfg(t0=0.0, dt_graph=0.5, dt_node=0.1)
    ref, _ = oracle.graph_run(spec, cfg, 40.0)
    assert rel_err(_results(sim.graph), ref) < 1e-6


def test_mobility_simulation_edge_dampings(oracle):
    """The mobility example with damped mobility coefficients (minimum, per-group damping) and edge NPIs."""
    graph = ps.mobility_example_graph()
    for e in graph._edges:
        grp = e.property.coefficients
        grp[0].minimum = 0.2 * grp[0].baseline
        grp.add_damping(mio.MobilityDamping(np.full(10, 0.5), t=3.0, level=0, type=0))
        dn = e.property.dynamic_npis_infected
        dn.duration = 5.0
        dn.base_value = 100000
        dn.set_threshold(100.0, [mio.DampingSampling(0.7, 1, 0, 0.0, [0], np.ones(1))])
    spec = _spec(graph)
    sim = osecir.MobilitySimulation(graph, 0.0, dt=0.5)
    sim.advance(30)
    cfg = capi.sim_cfg(t0=0.0, dt_graph=0.5, dt_node=0.1)
    ref, _ = oracle.graph_run(spec, cfg, 30.0)
    assert rel_err(_results(sim.graph), ref) < 1e-6


def test_ensemble_statistics_of_a_study(oracle):
    """osecir.interpolate_ensemble_results, ensemble_mean and ensemble_percentile over the graphs of a study
    (computed on the device from the host results, mgo_ensemble_*_values) equal the oracle's restatement of
    analyze_result.h:289-343 bitwise."""
    graph = ps.study_graph(n=4, G=2)
    study = osecir.GraphParameterStudy(graph, t0=0, tmax=12, dt=0.5, num_runs=7)
    graphs = study.run(sample=ps.scale_contacts)
    ens = osecir.interpolate_ensemble_results(graphs)
    assert len(ens) == 7 and len(ens[0]) == 4 and ens[0][0].get_num_time_points() == 13
    vals = [np.stack([np.stack(ts.values) for ts in run]) for run in ens]   # [run] -> [node][time][element]
    assert not np.array_equal(vals[0], vals[1])
    mean = osecir.ensemble_mean(ens)
    assert len(mean) == 4 and mean[2].get_time(12) == 12
    assert np.array_equal(np.stack([np.stack(ts.values) for ts in mean]), oracle.ensemble_mean(vals))
    for p in (0.1, 0.5, 0.9):
        pct = osecir.ensemble_percentile(ens, p)
        assert np.array_equal(np.stack([np.stack(ts.values) for ts in pct]), oracle.ensemble_percentile(vals, p)), p


def test_parameter_study_example(oracle):
    """The reference's ode_secir_parameter_study.py: every run samples TransmissionProbabilityOnContact from
    U(0.1, 0.2) (shared by both nodes, drawn per group) and its device result equals the oracle on the
    simulated graph."""
    graph = ps.parameter_study_example_graph()
    study = osecir.GraphParameterStudy(graph, t0=0, tmax=10, dt=0.5, num_runs=3)
    seen = []

    def handle_result(g, run_idx):
        tpc = [g.get_node(0).property.model.parameters.TransmissionProbabilityOnContact[mio.AgeGroup(i)].value
               for i in range(6)]
        assert all(0.1 <= v <= 0.2 for v in tpc)
        assert tpc == [g.get_node(1).property.model.parameters.TransmissionProbabilityOnContact[mio.AgeGroup(i)].value
                       for i in range(6)]
        assert g.get_node(0).property.result.get_last_time() == 10
        cfg = capi.sim_cfg(t0=0.0, dt_graph=0.5, dt_node=0.5)
        ref, _ = oracle.graph_run(_spec(g), cfg, 10.0)
        assert rel_err(_results(g), ref) < 1e-6
        seen.append(tuple(tpc))

    mio.seed_random_number_generator(3)
    study.run(handle_result)
    assert len(seen) == 3 and len(set(seen)) == 3


def test_mobility_parameter_study_example(oracle):
    """The reference's mobility_parameter_study.py: node 0's initial InfectedNoSymptoms is drawn per run from
    the truncated normal on [50, 2000], the susceptible node keeps none, the group total stays 10000; each
    run's device result equals the oracle on its graph."""
    graph = ps.mobility_parameter_study_graph()
    study = osecir.GraphParameterStudy(graph, 0, 50, dt=1.0, num_runs=3)
    carriers = []

    def handle_result(g, run_idx):
        m0, m1 = g.get_node(0).property.model, g.get_node(1).property.model
        ins = m0.populations[mio.AgeGroup(0), osecir.InfectionState.InfectedNoSymptoms].value
        assert 50 <= ins <= 2000
        assert m1.populations[mio.AgeGroup(0), osecir.InfectionState.InfectedNoSymptoms].value == 0
        assert abs(m0.populations.get_group_total_AgeGroup(0) - 10000) < 1e-9
        assert g.get_node(1).property.result.get_last_time() == 50
        cfg = capi.sim_cfg(t0=0.0, dt_graph=1.0, dt_node=1.0)
        ref, _ = oracle.graph_run(_spec(g), cfg, 50.0)
        assert rel_err(_results(g), ref) < 1e-6
        carriers.append(ins)

    mio.seed_random_number_generator(8)
    study.run(handle_result)
    assert len(carriers) == 3 and len(set(carriers)) == 3
