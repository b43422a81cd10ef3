"""The oracle against the reference's golden fixtures and known answers (CPU only).

These pin the oracle before it is trusted as the checker of the device kernels.
"""
import numpy as np
import pytest

from memilio_amd import capi, synthetic
from tests import setups


def _compare_series(times, vals, ref, tol, groups=1):
    assert len(times) == ref.shape[0], f"number of time points {len(times)} != {ref.shape[0]}"
    np.testing.assert_allclose(times, ref[:, 0], rtol=0, atol=tol)
    summed = vals.reshape(len(times), groups, 10).sum(axis=1) if groups > 1 else vals
    ncol = ref.shape[1] - 1
    np.testing.assert_allclose(summed[:, :ncol], ref[:, 1:], rtol=0, atol=tol)


def test_secihurd_rkf45_single(oracle):
    """test_odesecir.cpp:39-112 compareWithPreviousRun: RKF45, dt in [0.3, 1], rel 1e-3, abs 1e-1."""
    spec = setups.secihurd_single(cps=0.3, set_total_first=True)
    cfg = setups.rkf45_cfg(rel=1e-3, abs_=1e-1, dt_min=0.3, dt_max=1.0)
    t, v = oracle.simulate(spec, cfg, 50.0)
    _compare_series(t, v, setups.load_csv("secihurd-compare.csv"), 1e-10)


def test_secihurd_rkf45_ageres(oracle):
    """test_odesecir_ageres.cpp:27-108: G=3, RKF45, rel 1e-4, summed over age groups."""
    spec = setups.secihurd_ageres(cps=0.3)
    cfg = setups.rkf45_cfg(rel=1e-4, abs_=1e-1, dt_min=0.3, dt_max=1.0)
    t, v = oracle.simulate(spec, cfg, 50.0)
    _compare_series(t, v, setups.load_csv("secihurd-compare.csv"), 1e-10, groups=3)


def test_secihurd_cashkarp_ageres(oracle):
    """test_odesecir_ageres.cpp:110-189: G=3, odeint Cash-Karp, dt in [0.3, 1], rel 1e-4, abs 1e-1."""
    spec = setups.secihurd_ageres(cps=0.25)
    cfg = setups.ck_cfg(rel=1e-4, abs_=1e-1, dt_min=0.3, dt_max=1.0, dt=0.3)
    t, v = oracle.simulate(spec, cfg, 50.0)
    _compare_series(t, v, setups.load_csv("secihurd-compare-cashkarp.csv"), 1e-10, groups=3)


def test_default_integrator_single(oracle):
    """pycode test_osecir.py:122-150: default core (Cash-Karp, abs 1e-10, rel 1e-5), dt0 0.1, tmax 50.

    The time column pins the odeint controller: 0.1 -> 0.55 is the x4.5 growth cap and the rejection
    after t=25.3 the x0.2 shrink floor."""
    spec = setups.secihurd_single(cps=0.25, set_total_first=False)
    cfg = setups.ck_cfg()
    t, v = oracle.simulate(spec, cfg, 50.0)
    _compare_series(t, v, setups.load_csv("ode-secihurd-compare.csv"), 1e-10)


def test_default_integrator_ageres(oracle):
    """pycode test_osecir_ageres.py:116-150: G=3, default core, summed over age groups, 1e-10."""
    spec = setups.secihurd_ageres(cps=0.25)
    cfg = setups.ck_cfg()
    t, v = oracle.simulate(spec, cfg, 50.0)
    _compare_series(t, v, setups.load_csv("ode-secihurd-ageres-compare.csv"), 1e-10, groups=3)


def test_graph_fixture_oseir(oracle):
    """test_graph_simulation.cpp:278-339 consistencyFlowMobility: node 0 of the 3-node SEIR graph."""
    params, contact, pops, es, ee, coeffs = setups.oseir_fixture_inputs()
    t, v = oracle.oseir_graph_series(params, contact, pops, es, ee, coeffs, 0.0, 0.001, 0.1, 1.0, 0)
    ref = setups.load_csv("graphsimulation-compare.csv")
    assert len(t) == ref.shape[0]
    np.testing.assert_allclose(t, ref[:, 0], rtol=0, atol=1e-10)
    np.testing.assert_allclose(v, ref[:, 1:], rtol=0, atol=1e-10)


def test_edge_apply_mobility_known_answers(oracle):
    """test_mobility.cpp:154-215 edgeApplyMobility: exact departure values and conserved returns."""
    spec = setups.mobility_pair_model()
    out = oracle.edge_scenario(spec, 3.125, np.full(10, 0.1))
    n1, n2 = out[0]
    exp1 = np.array([990 - 99, 0, 0, 0, 10 - 1, 0, 0, 0, 0, 0], float)
    exp2 = np.array([990 + 99, 0, 0, 0, 10 + 1, 0, 0, 0, 0, 0], float)
    assert np.array_equal(n1, exp1)
    assert np.array_equal(n2, exp2)
    v = out[1][0]
    assert v.sum() == pytest.approx(1000, rel=4 * np.finfo(float).eps)
    assert v[0] < 990 and abs(v[0] - 990) < 50
    assert v[1] > 0 and abs(v[1]) < 50
    assert v[2] > 0 and abs(v[2]) < 5
    assert abs(v[4] - 10) < 5
    assert v[6] > 0 and abs(v[6]) < 5
    assert out[1][1].sum() == pytest.approx(1000, rel=4 * np.finfo(float).eps)
    assert out[2][0].sum() == pytest.approx(900, rel=4 * np.finfo(float).eps)
    assert out[2][1].sum() == pytest.approx(1100, rel=4 * np.finfo(float).eps)


@pytest.mark.parametrize("kind", [capi.INTEGRATOR_CASH_KARP54, capi.INTEGRATOR_RKF45])
def test_adaptive_sine(oracle, kind):
    """test_numericalIntegration.cpp:119-160: adaptive cores (abs = rel = 1e-7, dt in [1e-3, 1]) on
    y' = cos t, y = sin t; error metric sqrt(sum e^2)/n < 1e-7 as the reference computes it."""
    cfg = capi.integrator_cfg(kind, 1e-7, 1e-7, 1e-3, 1.0)
    t, y = oracle.integrate_test_problem(cfg, 0.0, 2 * np.pi, 2 * np.pi / 10)
    assert t[-1] == pytest.approx(2 * np.pi, abs=1e-12)
    err = np.sqrt(np.sum((y - np.sin(t)) ** 2)) / (len(t) - 1)
    assert err < 1e-7


@pytest.mark.parametrize("kind", [capi.INTEGRATOR_CASH_KARP54_FIXED, capi.INTEGRATOR_EULER])
def test_fixed_step_sine(oracle, kind):
    """test_numericalIntegration.cpp:82-109: fixed-step cores on y' = cos t with n = 1000 steps; error
    metric sqrt(sum e^2)/n < 1e-3; the 5th-order Cash-Karp solution is also checked far tighter."""
    cfg = capi.integrator_cfg(kind, 0, 0, 0, 0)
    t, y = oracle.integrate_test_problem(cfg, 0.0, 2 * np.pi, 2 * np.pi / 1000)
    assert len(t) == 1001
    err = np.sqrt(np.sum((y - np.sin(t)) ** 2)) / 1000
    assert err < 1e-3
    if kind == capi.INTEGRATOR_CASH_KARP54_FIXED:
        assert np.max(np.abs(y - np.sin(t))) < 1e-12


def test_adaptive_step_size_contract(oracle):
    """test_numericalIntegration.cpp:162-247 adaptiveStepSizing: every step lies in [dt_min, dt_max]."""
    for kind in (capi.INTEGRATOR_CASH_KARP54, capi.INTEGRATOR_RKF45):
        cfg = capi.integrator_cfg(kind, 1e-6, 1e-6, 0.01, 0.5)
        t, _ = oracle.integrate_test_problem(cfg, 0.0, 10.0, 0.1)
        dts = np.diff(t)
        assert np.all(dts[:-1] >= 0.01 - 1e-15) and np.all(dts <= 0.5 + 1e-15)
        assert t[-1] == pytest.approx(10.0, abs=1e-12)


def test_fixup_scenario_triggers(oracle):
    """The GPU fix-up tests rely on this set-up driving the return fix-up; keep it doing so."""
    spec = setups.fixup_graph()
    for cfg in (synthetic.fixed_cfg(0.1), synthetic.ck_cfg(1e-10)):
        _, st = oracle.graph_run(spec, cfg, 20.0)
        assert st.negative_fixups > 0
        assert st.returns > 0


def test_ensemble_analysis_known_answers(oracle):
    """The oracle's ensemble_mean / ensemble_percentile / sum_nodes against the known answers of the
    reference's TestEnsembleMean, TestEnsemblePercentile and TestSumNodes (cpp/tests/test_analyze_result.cpp:
    220-380), arrays laid out [run][time][node][element]."""
    # sum_nodes: 2 runs, 3 nodes, 3 times, 1 element
    r1 = np.array([[0, 3, 6], [1, 4, 7], [2, 5, 8]], float)[:, :, None]
    r2 = np.array([[0.5, 7.5, 1.5], [2.0, 9.5, 2.5], [5.0, 1.0, 3.0]])[:, :, None]
    s = oracle.sum_nodes([r1, r2])
    assert s.shape == (2, 3, 1)
    assert np.allclose(s[0, :, 0], [9, 12, 15]) and np.allclose(s[1, :, 0], [9.5, 14, 9])
    # ensemble_mean: 2 runs, 2 nodes, 3 times, 1 element
    m1 = np.array([[0, 0], [1, 1], [2, 2]], float)[:, :, None]
    m2 = np.array([[0.5, 1.5], [3.0, 0.5], [0.0, 1.0]])[:, :, None]
    mean = oracle.ensemble_mean([m1, m2])
    assert np.allclose(mean[:, 0, 0], [0.25, 2.0, 1.0]) and np.allclose(mean[:, 1, 0], [0.75, 0.75, 1.5])
    # ensemble_percentile: 4 runs, 2 nodes, 3 times, 2 elements
    runs = [np.zeros((3, 2, 2)) for _ in range(4)]
    runs[0][0, 1, 0] = 0.2
    runs[1][1, 0, 1], runs[1][0, 1, 0] = 1.0, 0.1
    runs[2][1, 0, 1], runs[2][0, 1, 0] = 2.0, 0.3
    runs[3][1, 0, 1] = 3.0
    for p, a, b in ((0.2, 0.0, 0.0), (0.4, 1.0, 0.1), (0.7, 2.0, 0.2), (0.9, 3.0, 0.3)):
        q = oracle.ensemble_percentile(runs, p)
        assert q[1, 0, 1] == a and q[0, 1, 0] == b, p


# ------------------------------------------------------------------------------------------ dynamic NPIs
# Known answers of cpp/tests/test_dynamic_npis.cpp (the scenarios restated as data).

def _damping_list(lst):
    return [(t, lev, typ, np.round(np.asarray(c, float), 12).tolist()) for (t, lev, typ, c) in lst]


def test_npi_implement_empty(oracle):
    """test_dynamic_npis.cpp:145-170: two empty vector-shaped lists, NPI 0.8 on both over [0.45, 0.6]."""
    out = oracle.implement_npis(2, 1, [[], []], [(0.8, 0, 0, [0, 1], [1.0, 1.0])], 0.45, 0.6, make_matrix=0)
    for lst in out:
        assert _damping_list(lst) == [(0.45, 0, 0, [0.8, 0.8]), (0.6, 0, 0, [0.0, 0.0])]


def test_npi_implement(oracle):
    """test_dynamic_npis.cpp:172-262: existing dampings before, during and after the NPI, other types and
    levels, a second NPI overlapping the first."""
    lists = [[(0.3, 0, 0, 0.15), (0.4, 0, 0, 0.2), (0.5, 0, 1, 0.25), (0.7, 0, 0, 0.35)],
             [(0.5, 1, 1, 0.25), (0.9, 0, 0, 0.45)]]
    lists = [[(t, lev, typ, [c] * 3) for (t, lev, typ, c) in lst] for lst in lists]
    out = oracle.implement_npis(3, 1, lists, [(0.4, 0, 0, [0, 1], [1.0, 1.0, 1.0])], 0.45, 0.6, make_matrix=0)
    v = lambda c: [c] * 3   # noqa: E731
    assert _damping_list(out[0]) == [(0.3, 0, 0, v(0.15)), (0.4, 0, 0, v(0.2)), (0.45, 0, 0, v(0.4)),
                                     (0.5, 0, 1, v(0.25)), (0.6, 0, 0, v(0.2)), (0.7, 0, 0, v(0.35))]
    assert _damping_list(out[1]) == [(0.45, 0, 0, v(0.4)), (0.5, 1, 1, v(0.25)), (0.6, 0, 0, v(0.0)),
                                     (0.9, 0, 0, v(0.45))]
    out2 = oracle.implement_npis(3, 1, out, [(0.3, 0, 0, [0], [1.0, 1.0, 1.0])], 0.3, 0.9, make_matrix=0)
    assert _damping_list(out2[0]) == [(0.3, 0, 0, v(0.3)), (0.45, 0, 0, v(0.4)), (0.5, 0, 1, v(0.25)),
                                      (0.6, 0, 0, v(0.3)), (0.7, 0, 0, v(0.35))]
    assert _damping_list(out2[1]) == _damping_list(out[1])


def _npi_model(isy, threshold, base, value=1.0, duration=5.0, delay=0.0, begin=None, end=None, contact=None):
    """osecir Model(1) of the secir_* tests: InfectedSymptoms = isy, Susceptible = 100 - isy."""
    from memilio_amd.graphspec import GraphSpec, ISY, S
    spec = GraphSpec(1, 1)
    spec.populations[0, 0, ISY] = isy
    spec.populations[0, 0, S] = 100.0 - isy
    spec.contact_baseline[0] = 1.0 if contact is None else contact
    cfg = {"thresholds": [(threshold, [(value, 0, 0, [1.0])])], "duration": duration, "delay": delay,
           "base_value": base}
    if begin is not None:
        cfg["directive_begin"] = begin
    if end is not None:
        cfg["directive_end"] = end
    spec.set_dynamic_npis([cfg])
    return spec


def test_npi_secir_threshold(oracle):
    """test_dynamic_npis.cpp:421-482: below the threshold nothing happens; above it two dampings appear."""
    _, n = oracle.npi_frozen_run(_npi_model(1.0, 0.05 * 23000, 23000), 0.0, 3.0, [0.0])
    assert n == 0
    _, n = oracle.npi_frozen_run(_npi_model(10.0, 0.05 * 50000, 50000), 0.0, 3.0, [0.0])
    assert n == 2


def test_npi_secir_delayed_implementation(oracle):
    """test_dynamic_npis.cpp:484-539."""
    c, _ = oracle.npi_frozen_run(_npi_model(10.0, 2500, 50000, 0.5, delay=3.0), 0.0, 3.0, [0.0, 1.0, 3.0])
    assert c.tolist() == [0.5, 0.5, 0.5]
    c, _ = oracle.npi_frozen_run(_npi_model(10.0, 2500, 50000, 0.5, delay=2.0), 1.0, 5.0, [3.0, 4.0])
    assert c.tolist() == [1.0, 0.5]
    c, _ = oracle.npi_frozen_run(_npi_model(10.0, 2500, 50000, 0.5, delay=10.0), 1.0, 4.0, [11.0, 12.0])
    assert c.tolist() == [1.0, 0.5]


def test_npi_secir_directives(oracle):
    """test_dynamic_npis.cpp:541-628."""
    m = lambda **kw: _npi_model(10.0, 2500, 50000, 0.5, **kw)   # noqa: E731
    c, _ = oracle.npi_frozen_run(m(begin=5.0), 0.0, 3.0, [0.0, 1.0, 3.0])
    assert c.tolist() == [1.0, 1.0, 1.0]
    c, _ = oracle.npi_frozen_run(m(delay=2.0, begin=0.0), 0.0, 3.0, [0.0, 4.0, 6.0])
    assert c.tolist() == [0.5, 0.5, 1.0]
    c, _ = oracle.npi_frozen_run(m(delay=2.0, begin=0.0, end=3.0), 0.0, 5.0, [0.0, 2.0, 4.0])
    assert c.tolist() == [0.5, 0.5, 1.0]
    c, _ = oracle.npi_frozen_run(m(delay=2.0, begin=0.0, end=1e6), 1.0, 5.0, [3.0, 4.0, 8.0, 9.0])
    assert c.tolist() == [1.0, 0.5, 0.5, 1.0]
    c, _ = oracle.npi_frozen_run(m(delay=2.0, begin=0.0, end=4.0), 1.0, 6.0, [1.0, 4.0, 5.0])
    assert c.tolist() == [1.0, 0.5, 1.0]


def test_npi_secir_backward_consistency_and_keep_alive(oracle):
    """test_dynamic_npis.cpp:630-722: the dynamic NPI equals predefined dampings at t_start+1 / t_end+1;
    an NPI that expires while the threshold is still exceeded is renewed without a dip."""
    ts = [1.0, 2.0, 3.0, 3.5, 4.0, 5.0, 7.0, 8.0, 8.5, 9.0]
    c, _ = oracle.npi_frozen_run(_npi_model(10.0, 2500, 50000, 0.5, delay=2.0), 1.0, 8.0, ts)
    pre = _npi_model(10.0, 2500, 50000, 0.5, delay=2.0)
    pre.set_dampings([[(4.0, 0, 0, 0.5), (9.0, 0, 0, 0.0)]])
    pre.npis = None
    pre._desc = None
    expect = [oracle.contact_at(pre, t, 0, 0) for t in ts]
    assert c.tolist() == expect
    c, _ = oracle.npi_frozen_run(_npi_model(10.0, 2500, 50000, 0.5, delay=2.0), 1.0, 12.0,
                                 [4.0, 8.0, 9.0, 10.0, 13.0, 14.0])
    assert c.tolist() == [0.5, 0.5, 0.5, 0.5, 0.5, 1.0]


def test_npi_scenario_triggers(oracle):
    """The graph scenario of the device NPI tests implements NPIs (guards against a vacuous comparison)."""
    for kw in (dict(duration=14.0, delay=3.0), dict(duration=2.3, delay=0.7, directive_begin=2.0,
                                                      directive_end=40.2)):
        g = synthetic.with_npis(synthetic.graph(num_nodes=30, G=3, seed=61, mean_degree=6.0), **kw)
        g0 = synthetic.graph(num_nodes=30, G=3, seed=61, mean_degree=6.0)
        a, _ = oracle.graph_run(g, synthetic.fixed_cfg(0.1), 60.0)
        b, _ = oracle.graph_run(g0, synthetic.fixed_cfg(0.1), 60.0)
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6)) > 1e-2


def test_npi_mobility_edge(oracle):
    """test_dynamic_npis.cpp:319-376: the edge's NPI check at each apply_mobility call (source state (0, 1)
    at t = 0.5, (0.01, 0.99) at 3, (0.02, 0.98) at 4.5 and 6; threshold 1500 per 100,000, duration 5)."""
    counts = oracle.edge_npi_calls(2, 0.015 * 100000, 100000, 5.0, 1.0, [1.0, 1.0], [0.5, 3.0, 4.5, 6.0],
                                   [0.0, 0.01, 0.02, 0.02])
    assert counts == [0, 0, 2, 2]


def test_mobility_dampings_scenario_triggers(oracle):
    """The graph scenario of the device edge-damping tests: out-edge dampings and edge NPIs change the run."""
    g = synthetic.with_edge_dampings(synthetic.graph(num_nodes=30, G=3, seed=64, mean_degree=6.0))
    g0 = synthetic.graph(num_nodes=30, G=3, seed=64, mean_degree=6.0)
    a, _ = oracle.graph_run(g, synthetic.fixed_cfg(0.1), 40.0)
    b, _ = oracle.graph_run(g0, synthetic.fixed_cfg(0.1), 40.0)
    assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6)) > 1e-3


# ------------------------------------------------------------------ mobility factors and commuter testing
def _departure(oracle, spec, t, c=0.1):
    """(source, destination) after one departure of a G-group node pair with all coefficients c."""
    out = oracle.edge_scenario(spec, t, np.full(spec.state_size(), c))
    return out[0][0], out[0][1]


def _factor_model(G=1, ttc=45.0, beta=(0.25,), max_beta=(0.5,), ins=(100.0,)):
    """test_odesecir.cpp:1052-1063: T_E 3, T_INS 2, RecoveredPerInfectedNoSymptoms 0.1, INS 100 (split over
    the groups); 40 symptomatic and 1000 susceptible per group added so that the ISy factor shows in the
    travelers (the factors depend on INS only)."""
    from memilio_amd.graphspec import GraphSpec, S, INS, ISY
    spec = GraphSpec(1, G)
    a = spec.age_params[0]
    a[capi.A_TIME_EXPOSED] = 3.0
    a[capi.A_TIME_INFECTED_NO_SYMPTOMS] = 2.0
    a[capi.A_RECOVERED_PER_INFECTED_NO_SYMPTOMS] = 0.1
    a[capi.A_RISK_OF_INFECTION_FROM_SYMPTOMATIC] = beta
    a[capi.A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC] = max_beta
    spec.node_scalars[0, capi.P_TEST_AND_TRACE_CAPACITY] = ttc
    spec.populations[0, :, INS] = ins
    spec.populations[0, :, ISY] = 40.0
    spec.populations[0, :, S] = 1000.0
    return spec


def _isy_factors(oracle, spec):
    from memilio_amd.graphspec import ISY, NC
    src, _ = _departure(oracle, spec, 0.0)
    isy = src.reshape(spec.G, NC)[:, ISY]
    return (40.0 - isy) / (0.1 * 40.0)


def test_mobility_factors_known_answers(oracle):
    """get_mobility_factors (test_odesecir.cpp:1052-1088) seen through a departure: test-and-trace load
    (1 - 0.1) / 2 * 100 = 45; capacity 45 gives the symptomatic factor beta, capacity 9 (x 5 = 45) gives
    max beta, capacity 20 a value strictly between; every other compartment travels with factor 1."""
    f = _isy_factors(oracle, _factor_model(ttc=45.0))
    assert f[0] == pytest.approx(0.25, rel=1e-12)
    f = _isy_factors(oracle, _factor_model(ttc=9.0))
    assert f[0] == pytest.approx(0.5, rel=1e-12)
    f_mid = _isy_factors(oracle, _factor_model(ttc=20.0))
    assert 0.25 < f_mid[0] < 0.5
    from memilio_amd.graphspec import INS, ISY, S, NC
    spec = _factor_model(ttc=20.0)
    src, dst = _departure(oracle, spec, 0.0)
    for k in (S, INS):
        assert src[k] == pytest.approx(spec.populations[0, 0, k] * 0.9, rel=1e-15)
        assert dst[k] == pytest.approx(spec.populations[0, 0, k] * 1.1, rel=1e-15)
    assert src.sum() + dst.sum() == pytest.approx(2 * spec.populations[0].sum(), rel=1e-15)
    # more groups: the load is the total over the groups, the factor each group's own
    f2 = _isy_factors(oracle, _factor_model(G=2, ttc=20.0, beta=(0.25, 0.25), max_beta=(0.5, 0.5),
                                            ins=(60.0, 40.0)))
    assert f2 == pytest.approx([f_mid[0]] * 2, rel=1e-12)
    f2 = _isy_factors(oracle, _factor_model(G=2, ttc=45.0, beta=(0.25, 0.1), max_beta=(0.5, 0.7),
                                            ins=(60.0, 40.0)))
    assert f2 == pytest.approx([0.25, 0.1], rel=1e-12)
    f2 = _isy_factors(oracle, _factor_model(G=2, ttc=9.0, beta=(0.25, 0.1), max_beta=(0.5, 0.7),
                                            ins=(60.0, 40.0)))
    assert f2 == pytest.approx([0.5, 0.7], rel=1e-12)


@pytest.mark.parametrize("t,nondet", [(3.125, 0.4), (25.0, 1.0)], ids=["inside", "outside"])
def test_commuter_testing_known_answers(oracle, t, nondet):
    """test_commuters (test_odesecir.cpp:1086-1117) on departure: inside the detection window [0, 20) a
    fraction 1 - 0.4 of the departing INS and ISy travelers is detected and stays home as INSC / ISyC; the
    travelers keep 0.4 of them. Outside the window nothing is detected."""
    from memilio_amd.graphspec import GraphSpec, S, INS, INSC, ISY, ISYC, NC
    spec = GraphSpec(1, 2)
    spec.node_scalars[0, capi.P_COMMUTER_NONDETECTION] = 0.4
    spec.node_scalars[0, capi.P_START_COMMUTER_DETECTION] = 0.0
    spec.node_scalars[0, capi.P_END_COMMUTER_DETECTION] = 20.0
    spec.populations[0] = np.array([[800.0, 30.0, 50.0, 20.0, 40.0, 10.0, 5.0, 2.0, 100.0, 1.0],
                                    [600.0, 20.0, 70.0, 15.0, 60.0, 25.0, 8.0, 3.0, 80.0, 2.0]])
    c = 0.1
    y = spec.populations[0]
    src, dst = _departure(oracle, spec, t, c)
    src, dst = src.reshape(2, NC), dst.reshape(2, NC)
    for g in range(2):
        mob = c * y[g]       # default parameters: every mobility factor is 1
        assert dst[g, INS] == pytest.approx(y[g, INS] + mob[INS] * nondet, rel=1e-14)
        assert dst[g, ISY] == pytest.approx(y[g, ISY] + mob[ISY] * nondet, rel=1e-14)
        assert src[g, INSC] == pytest.approx(y[g, INSC] + mob[INS] * (1 - nondet) - mob[INSC], rel=1e-14, abs=1e-12)
        assert src[g, ISYC] == pytest.approx(y[g, ISYC] + mob[ISY] * (1 - nondet) - mob[ISYC], rel=1e-14, abs=1e-12)
        assert src[g, INS] == pytest.approx(y[g, INS] - mob[INS], rel=1e-14)
        assert src[g, S] == pytest.approx(y[g, S] - mob[S], rel=1e-15)
    assert src.sum() + dst.sum() == pytest.approx(2 * y.sum(), rel=1e-15)


def test_return_fixup_properties(oracle):
    """Return fix-up (metapopulation_mobility_instant.h:576-596), no known answers in the reference's tests:
    on a graph where it triggers, every node keeps its total at every output day (the supernumerous returns
    are moved inside the traveler vector, not created) and no compartment is negative afterwards."""
    spec = setups.fixup_graph()
    for cfg in (synthetic.fixed_cfg(0.1), synthetic.ck_cfg(1e-10)):
        got, st = oracle.graph_run(spec, cfg, 20.0)
        assert st.negative_fixups > 0
        per_node = got.sum(axis=2)
        assert np.max(np.abs(per_node - per_node[0]) / per_node[0]) < 1e-12
        assert got.min() >= -1e-9


@pytest.mark.parametrize("case", ["single", "tie", "groups"])
def test_return_fixup_known_answers(oracle, case):
    """Return fix-up (metapopulation_mobility_instant.h:576-596) against states derived by hand (see
    setups.fixup_known_answer): the deficit moves to the largest remaining compartment of the same age group,
    the first of equal maxima, with the remaining values taken once before the loop. Exact: every value of
    these set-ups is a short binary fraction."""
    spec, cfg, tmax, expected = setups.fixup_known_answer(case)
    got, st = oracle.graph_run(spec, cfg, tmax)
    assert got.shape == (2, 2, spec.G * 10)
    assert st.negative_fixups == (2 if case == "groups" else 1)
    np.testing.assert_array_equal(got[0].reshape(2, spec.G, 10), spec.populations)
    np.testing.assert_array_equal(got[1].reshape(2, spec.G, 10), expected)


def test_node_parallel_graph_run_is_bitwise_serial(oracle):
    """The oracle's graph run with its node phase on several OpenMP threads (the CPU timing of one graph
    parallel over nodes) equals the serial run bit for bit: nodes advance independently, edges stay in order."""
    spec = setups.fixup_graph()
    cfg = synthetic.ck_cfg(1e-10)
    a, sa = oracle.graph_run(spec, cfg, 12.0)
    b, sb = oracle.graph_run_threads(spec, cfg, 12.0, 4)
    assert np.array_equal(a, b)
    assert (sa.steps, sa.negative_fixups) == (sb.steps, sb.negative_fixups)
