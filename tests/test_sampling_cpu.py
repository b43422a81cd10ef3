"""Parameter sampling of the Python mirror: distributions, UncertainValue, UncertainContactMatrix and
osecir.draw_sample on a model and on a graph, after the reference's tests
cpp/tests/test_parameter_studies.cpp:34-325 and cpp/tests/test_uncertain.cpp:30-160. The generator is
numpy's, so the samples follow the reference's distributions but not its random values."""
import copy

import numpy as np
import pytest

import memilio_amd.simulation as mio
import memilio_amd.simulation.osecir as osecir

S = osecir.InfectionState
Q = 2.5758


# ----------------------------------------------------------------------------------------- distributions
def test_normal_distribution_adapts_standard_dev():
    """test_parameter_studies.cpp:187-237: a standard deviation too wide for [lb, ub] shrinks so that mean +-
    quantile * std fits; one that fits is kept; samples stay inside the bounds."""
    d = mio.ParameterDistributionNormal(-1.0, 1.0, 0.0, 1.0, Q)
    assert d.standard_dev == pytest.approx(1.0 / Q, rel=1e-15)
    d2 = mio.ParameterDistributionNormal(lb=-1.0, ub=1.0, mean=0.0, std_dev=d.standard_dev, quantile=Q)
    assert (d2.lower_bound, d2.upper_bound, d2.mean, d2.standard_dev) == \
        (d.lower_bound, d.upper_bound, d.mean, d.standard_dev)
    d3 = mio.ParameterDistributionNormal(1.0, 10.0, 5.0, 1.5, Q)
    assert d3.standard_dev == 1.5
    mio.seed_random_number_generator(11)
    vals = np.array([d3.get_sample() for _ in range(1000)])
    assert np.all(vals >= 1.0) and np.all(vals <= 10.0)
    # the distribution is the normal one: mean and spread of the samples (truncation at 2.67 / 3.33 sigma
    # changes the moments by well under the tolerance)
    assert abs(vals.mean() - 5.0) < 0.15 and abs(vals.std() - 1.5) < 0.15


def test_normal_distribution_mean_outside_bounds():
    """A mean outside [lb, ub] becomes (ub - lb) / 2 before the spread is fitted (parameter_distributions.h:
    220-245): for [2, 10] that is 4, not the midpoint 6."""
    d = mio.ParameterDistributionNormal(2.0, 10.0, 20.0, 1.0, Q)
    assert d.mean == 4.0
    assert d.standard_dev == pytest.approx(min(1.0, (4.0 - 2.0) / Q), rel=1e-15)


def test_normal_distribution_four_argument_form():
    """(lb, ub, mean, quantile): the standard deviation is the widest for which mean +- quantile * std fits."""
    d = mio.ParameterDistributionNormal(1.0, 5.0, 2.0, 2.0)
    assert d.standard_dev == pytest.approx(0.5, rel=1e-15)   # min((5 - 2) / 2, (2 - 1) / 2)
    d = mio.ParameterDistributionNormal(1.0, 5.0, 4.0, 2.0)
    assert d.standard_dev == pytest.approx(0.5, rel=1e-15)   # min((5 - 4) / 2, (4 - 1) / 2)


def test_normal_distribution_degenerate_bounds():
    """test_parameter_studies.cpp:231-236: a spread far below the interval's width, an interval so narrow that
    ten redraws miss it, and lb == ub."""
    mio.seed_random_number_generator(3)
    d = mio.ParameterDistributionNormal(0.999999999 * 3.0, 1.000000001 * 3.0, 3.0, 0.00000001)
    # four-argument form: 1e-8 is the quantile, so the spread is (ub - mean) / 1e-8 = 0.3, far wider than
    # the interval: the draws miss and the sample is clamped to a bound
    assert d.standard_dev == pytest.approx(0.3, rel=1e-6)
    v = d.get_sample()
    assert abs(v - 3.0) < 1e-7 and v in (d.lower_bound, d.upper_bound)
    assert mio.ParameterDistributionNormal(2.0, 2.0, 2.0, 1.0, Q).get_sample() == 2.0


def test_normal_distribution_retry_then_clamp():
    """The loop of parameter_distributions.h:332-347 with a scripted generator: redraws while outside the
    bounds, at most 10 times; after the tenth the value becomes ub if above, else lb."""
    class Gen:
        def __init__(self, seq):
            self.seq = list(seq)

        def normal(self, mean, std):
            return self.seq.pop(0)

    d = mio.ParameterDistributionNormal(0.0, 1.0, 0.5, 0.1, Q)
    assert d.get_rand_sample(Gen([0.3])) == 0.3
    assert d.get_rand_sample(Gen([7.0, -1.0, 0.9])) == 0.9
    assert d.get_rand_sample(Gen([7.0] * 11)) == 1.0
    assert d.get_rand_sample(Gen([-7.0] * 11)) == 0.0
    # the tenth redraw inside the bounds is still replaced by lb (the reference's loop does so)
    assert d.get_rand_sample(Gen([7.0] * 10 + [0.42])) == 0.0
    # the ninth redraw inside: kept
    assert d.get_rand_sample(Gen([7.0] * 9 + [0.42])) == 0.42


def test_uniform_distribution():
    """test_parameter_studies.cpp:239-253."""
    d = mio.ParameterDistributionUniform(1.0, 10.0)
    assert (d.lower_bound, d.upper_bound) == (1.0, 10.0)
    mio.seed_random_number_generator(5)
    vals = np.array([d.get_sample() for _ in range(1000)])
    assert np.all(vals >= 1.0) and np.all(vals <= 10.0)
    assert abs(vals.mean() - 5.5) < 0.3


def test_constant_distribution():
    """test_parameter_studies.cpp:272-278."""
    d = mio.ParameterDistributionConstant(2.0)
    assert d.constant == 2.0 and d.get_sample() == 2.0


def test_predefined_samples():
    """test_parameter_studies.cpp:280-325: predefined samples come first (even outside the bounds), once."""
    mio.seed_random_number_generator(9)
    u = mio.ParameterDistributionUniform(1.0, 10.0)
    u.add_predefined_sample(2)
    assert u.get_sample() == 2
    assert u.get_sample() != 2
    n = mio.ParameterDistributionNormal(-1.0, 1.0, 0.0, 0.1)
    n.add_predefined_sample(2)
    assert n.get_sample() == 2
    assert n.get_sample() != 2
    c = mio.ParameterDistributionConstant(3.0)
    c.add_predefined_sample(1.0)
    assert c.get_sample() == 1.0 and c.get_sample() == 3.0
    u.add_predefined_sample(4)
    u.add_predefined_sample(5)
    u.remove_predefined_samples()
    assert u.get_predefined_samples() == []


def test_seeded_generator_repeats():
    d = mio.ParameterDistributionNormal(0.0, 10.0, 5.0, 1.0, Q)
    mio.seed_random_number_generator(1, 2, 3)
    a = [d.get_sample() for _ in range(5)]
    mio.seed_random_number_generator(1, 2, 3)
    b = [d.get_sample() for _ in range(5)]
    assert a == b and len(set(a)) == 5


def test_distribution_json():
    """parameter_distributions.h:363-432, 505-559, 867-917: Type, the fields and PredefinedSamples."""
    n = mio.ParameterDistributionNormal(1.0, 10.0, 5.0, 1.5, Q)
    n.add_predefined_sample(7.0)
    js = n.to_json()
    assert js == {"Type": "Normal", "Mean": 5.0, "StandardDev": 1.5, "LowerBound": 1.0, "UpperBound": 10.0,
                  "Quantile": Q, "PredefinedSamples": [7.0]}
    back = mio.distribution_from_json(js)
    assert isinstance(back, mio.ParameterDistributionNormal) and back.to_json() == js
    for d in (mio.ParameterDistributionUniform(1.0, 2.0), mio.ParameterDistributionConstant(4.0)):
        assert mio.distribution_from_json(d.to_json()).to_json() == d.to_json()
    ln = {"Type": "LogNormal", "LogMean": 0.0, "LogStddev": 0.25, "PredefinedSamples": [-5.0]}
    d = mio.distribution_from_json(ln)
    assert isinstance(d, mio.ParameterDistributionLogNormal) and d.to_json() == ln
    assert d.get_sample() == -5.0 and d.get_sample() > 0
    ex = {"Type": "Exponential", "Rate": 2.0, "PredefinedSamples": []}
    assert mio.distribution_from_json(ex).to_json() == ex
    with pytest.raises(ValueError, match="not valid"):
        mio.distribution_from_json({"Type": "Poisson", "PredefinedSamples": []})
    with pytest.raises(KeyError):
        mio.distribution_from_json({"Type": "Uniform", "LowerBound": 0.0, "PredefinedSamples": []})
    # the reader uses the five-argument constructor: a written spread too wide for the bounds is adapted
    wide = {"Type": "Normal", "Mean": 5.0, "StandardDev": 10.0, "LowerBound": 1.0, "UpperBound": 10.0,
            "Quantile": Q, "PredefinedSamples": []}
    assert mio.distribution_from_json(wide).standard_dev == pytest.approx(4.0 / Q, rel=1e-15)


# ----------------------------------------------------------------------------------------- uncertain values
def _normal_around(v, dev_rel=0.2):
    return mio.ParameterDistributionNormal(max(1e-6, (1 - dev_rel * 2.6) * v), (1 + dev_rel * 2.6) * v, v,
                                           dev_rel * v)


def test_uncertain_value_basic_and_copy():
    """test_uncertain.cpp:30-52: value access; a copy has its own, equal distribution."""
    val = mio.UncertainValue(3.0)
    assert val.value == 3.0
    val.value = 2.0
    assert float(val) == 2.0
    val.set_distribution(_normal_around(2.0))
    val2 = copy.deepcopy(val)
    assert val2.value == 2.0
    assert val2.get_distribution() is not val.get_distribution()
    assert val2.get_distribution().to_json() == val.get_distribution().to_json()


def test_uncertain_value_random_sample():
    """test_uncertain.cpp:54-72 with a scripted distribution."""
    class Scripted(mio.ParameterDistribution):
        def __init__(self, seq):
            super().__init__()
            self.seq = list(seq)

        def get_rand_sample(self, gen):
            return self.seq.pop(0)

    val = mio.UncertainValue(2.0)
    val.set_distribution(Scripted([0.5, 1.5]))
    assert val.value == 2.0
    val.draw_sample()
    assert val.value == 0.5
    val.value = 1.0
    assert val.value == 1.0
    val.draw_sample()
    assert val.value == 1.5


def test_uncertain_value_without_distribution_keeps_value():
    val = mio.UncertainValue(4.0)
    assert val.draw_sample() == 4.0 and val.value == 4.0


def test_uncertain_value_predef():
    """test_uncertain.cpp:96-114: the predefined sample is copied with the distribution."""
    mio.seed_random_number_generator(2)
    val = mio.UncertainValue(3.0)
    val.set_distribution(_normal_around(3.0))
    val.get_distribution().add_predefined_sample(17111.0)
    val2 = copy.deepcopy(val)
    val.draw_sample()
    assert val.value == 17111.0
    val.draw_sample()
    assert val.value != 17111.0
    val2.draw_sample()
    assert val2.value == 17111.0


def test_uncertain_matrix():
    """test_uncertain.cpp:116-160: sampled dampings and the school-holiday damping (on at a holiday's first
    day, off at its last), accumulated or replacing the matrix's dampings."""
    base = np.array([[1.0, 2.0], [2.0, 4.0]])
    cm = mio.ContactMatrix(base)
    cm.add_damping(mio.Damping(np.full((2, 2), 0.7), t=30.0))
    grp = mio.ContactMatrixGroup(1, 2)
    grp[0] = cm
    um = mio.UncertainContactMatrix(grp)
    assert um.cont_freq_mat[0].baseline[0, 1] == 2 and um.cont_freq_mat[0].baseline[1, 1] == 4
    assert um.cont_freq_mat[0].dampings[0].coeffs[1, 1] == 0.7
    um.dampings.append(mio.DampingSampling(0.5, 0, 0, 3.0, [0], np.ones(2)))
    um.school_holiday_damping = mio.DampingSampling(1.0, 1, 0, 0.0, [0], np.array([1.0, 0.0]))
    um.school_holidays = [(5.0, 17.0)]
    um.draw_sample(True)
    assert len(um.cont_freq_mat[0].dampings) == 4
    um.draw_sample()
    ds = um.cont_freq_mat[0].dampings
    assert len(ds) == 3
    assert (ds[0].level, ds[0].coeffs[0, 0], ds[0].time) == (0, 0.5, 3.0)
    assert (ds[1].level, ds[1].coeffs[0, 0], ds[1].time) == (1, 1.0, 5.0)
    assert (ds[2].level, ds[2].coeffs[0, 0], ds[2].time) == (1, 0.0, 17.0)
    # make_contact_damping_matrix: d_ij = 1 - sqrt((1 - g_i)(1 - g_j)) with g = 1.0 * (1, 0)
    np.testing.assert_array_equal(ds[1].coeffs, [[1.0, 1.0], [1.0, 0.0]])
    # the holiday damping keeps its own value after make_matrix
    assert um.school_holiday_damping.value.value == 1.0


def test_damping_matrices():
    """make_contact_damping_matrix against its closed form; make_mobility_damping_vector repeats each group's
    value over the group's compartments (damping_sampling.h:268-306)."""
    g = np.array([0.2, 0.5, 0.0])
    m = mio.make_contact_damping_matrix(g)
    for i in range(3):
        for j in range(3):
            assert m[i, j] == pytest.approx(1 - np.sqrt((1 - g[i]) * (1 - g[j])), abs=1e-15)
    np.testing.assert_allclose(np.diag(m), g, atol=1e-15)
    np.testing.assert_array_equal(mio.make_mobility_damping_vector(30, g), np.repeat(g, 10))


# ----------------------------------------------------------------------------------------- draw_sample
def _secir_model(G=3):
    """The model of test_parameter_studies.cpp:34-95 (sample_from_secir_params, sample_graph)."""
    model = osecir.Model(G)
    fact = 1.0 / G
    p = model.parameters
    for g in range(G):
        A = mio.AgeGroup(g)
        p.TimeExposed[A] = 2.6
        p.TimeInfectedNoSymptoms[A] = 2.6
        p.TimeInfectedSymptoms[A] = 5.0
        p.TimeInfectedSevere[A] = 10.0
        p.TimeInfectedCritical[A] = 8.0
        p.Seasonality = 0.0
        p.ICUCapacity = 100.0
        p.TestAndTraceCapacity = 10.0
        pop = model.populations
        pop[A, S.Exposed] = fact * 100
        pop[A, S.InfectedNoSymptoms] = fact * 50
        pop[A, S.InfectedSymptoms] = fact * 50
        pop[A, S.InfectedSevere] = fact * 20
        pop[A, S.InfectedCritical] = fact * 10
        pop[A, S.Recovered] = fact * 10
        pop[A, S.Dead] = 0.0
        pop.set_difference_from_group_total_AgeGroup((A, S.Susceptible), fact * 10000)
        p.TransmissionProbabilityOnContact[A] = 0.05
        p.RelativeTransmissionNoSymptoms[A] = 0.67
        p.RecoveredPerInfectedNoSymptoms[A] = 0.09
        p.RiskOfInfectionFromSymptomatic[A] = 0.25
        p.MaxRiskOfInfectionFromSymptomatic[A] = 0.85
        p.SeverePerInfectedSymptoms[A] = 0.2
        p.CriticalPerSevere[A] = 0.25
        p.DeathsPerCritical[A] = 0.3
    p.ContactPatterns.cont_freq_mat[0] = mio.ContactMatrix(np.full((G, G), fact * 10.0))
    return model


def test_set_params_distributions_normal():
    """parameter_space.h:46-118: bounds (1 -+ 2.6 dev_rel) * value (at least 0.001, 0 for seasonality and
    compartments), a constant where the bounds meet, and one sampled contact damping at mid-time."""
    model = _secir_model()
    osecir.set_params_distributions_normal(model, 0.0, 100.0, 0.2)
    p = model.parameters
    d = p.TimeInfectedSevere[1].get_distribution()
    assert isinstance(d, mio.ParameterDistributionNormal)
    assert d.lower_bound == pytest.approx(0.48 * 10.0, rel=1e-15)
    assert d.upper_bound == pytest.approx(1.52 * 10.0, rel=1e-15)
    assert d.mean == 10.0 and d.quantile == pytest.approx(0.2 * 10.0, rel=1e-15)
    # the four-argument form: the widest spread that fits with the given quantile
    assert d.standard_dev == pytest.approx(min(d.upper_bound - 10.0, 10.0 - d.lower_bound) / d.quantile, rel=1e-12)
    # seasonality 0: both bounds 0 -> a constant distribution
    s = p.Seasonality.get_distribution()
    assert isinstance(s, mio.ParameterDistributionConstant) and s.constant == 0.0
    e = model.populations[mio.AgeGroup(2), S.Exposed].get_distribution()
    assert e.lower_bound == pytest.approx(0.48 * 100 / 3, rel=1e-14)
    assert model.populations[mio.AgeGroup(0), S.Susceptible].get_distribution() is None
    assert model.populations[mio.AgeGroup(0), S.Dead].get_distribution() is None
    cp = p.ContactPatterns
    assert len(cp.dampings) == 1
    ds = cp.dampings[0]
    assert (ds.time, ds.level, ds.type, ds.matrix_indices) == (50.0, 0, 0, [0])
    assert ds.value.value == 0.5 and isinstance(ds.value.get_distribution(), mio.ParameterDistributionNormal)


def _in_bounds(uv):
    d = uv.get_distribution()
    if isinstance(d, mio.ParameterDistributionNormal):
        return d.lower_bound <= uv.value <= d.upper_bound
    return True


def test_sample_from_secir_params():
    """test_parameter_studies.cpp:34-108: the shared parameters are equal across groups, the group totals are
    kept, and the contact matrix carries the one sampled damping."""
    mio.seed_random_number_generator(21)
    model = _secir_model()
    osecir.set_params_distributions_normal(model, 0.0, 100.0, 0.2)
    model.parameters.ContactPatterns.cont_freq_mat.add_damping(mio.Damping(np.full((3, 3), 0.3), t=5.0))
    ref = copy.deepcopy(model)
    osecir.draw_sample(model)
    p = model.parameters
    for name in ("TimeExposed", "TimeInfectedNoSymptoms", "RelativeTransmissionNoSymptoms",
                 "RiskOfInfectionFromSymptomatic", "MaxRiskOfInfectionFromSymptomatic"):
        vals = [getattr(p, name)[i].value for i in range(3)]
        assert vals[0] == vals[1] == vals[2], name
        assert vals[0] != getattr(ref.parameters, name)[0].value, name
        assert _in_bounds(getattr(p, name)[0]), name
    for i in range(3):
        A = mio.AgeGroup(i)
        assert model.populations.get_group_total_AgeGroup(A) == pytest.approx(10000 / 3, abs=1e-6)
        assert model.populations[A, S.Susceptible].value >= 0
        assert p.TransmissionProbabilityOnContact[A].value >= 0
        for name in ("TimeInfectedSymptoms", "TimeInfectedSevere", "TimeInfectedCritical", "DeathsPerCritical",
                     "TransmissionProbabilityOnContact", "SeverePerInfectedSymptoms", "CriticalPerSevere"):
            assert _in_bounds(getattr(p, name)[A]), (name, i)
        for st in (S.Exposed, S.InfectedNoSymptoms, S.InfectedSymptoms, S.InfectedSevere, S.InfectedCritical,
                   S.Recovered):
            assert _in_bounds(model.populations[A, st]), (st, i)
    # the age-specific parameters are drawn per group
    assert len({p.TimeInfectedSevere[i].value for i in range(3)}) == 3
    # the manual damping is replaced by the sampled one (make_matrix without accumulation)
    ds = p.ContactPatterns.cont_freq_mat[0].dampings
    assert len(ds) == 1 and ds[0].time == 50.0
    v = p.ContactPatterns.dampings[0].value.value
    np.testing.assert_allclose(ds[0].coeffs, np.full((3, 3), v), rtol=1e-14)
    # distributions are kept through the sample
    assert p.TimeExposed[2].get_distribution().to_json() == ref.parameters.TimeExposed[2].get_distribution().to_json()


def test_draw_sample_applies_constraints():
    """draw_sample ends with apply_constraints (parameter_space.h:203-209): a sampled probability above 1 is
    set to 0."""
    model = _secir_model()
    d = mio.ParameterDistributionUniform(1.5, 2.0)
    model.parameters.CriticalPerSevere[1].set_distribution(d)
    osecir.draw_sample(model)
    assert model.parameters.CriticalPerSevere[1].value == 0.0
    assert model.parameters.CriticalPerSevere[0].value == 0.25


def _sample_graph_setup():
    model = _secir_model()
    osecir.set_params_distributions_normal(model, 0.0, 100.0, 0.2)
    graph = osecir.ModelGraph()
    m0 = copy.deepcopy(model)
    m1 = copy.deepcopy(model)
    m1.populations[mio.AgeGroup(0), S.Exposed] = 80.0
    m1.populations.set_difference_from_group_total_AgeGroup((mio.AgeGroup(0), S.Susceptible), 5000.0)
    m1.parameters.ICUCapacity = 40.0
    m1.parameters.ICUCapacity.set_distribution(mio.ParameterDistributionUniform(30.0, 50.0))
    m1.parameters.ContactPatterns.school_holidays = [(10.0, 20.0)]
    m1.parameters.ContactPatterns.cont_freq_mat.add_damping(mio.Damping(np.full((3, 3), 0.9), t=1.0))
    graph.add_node(0, m0)
    graph.add_node(1, m1)
    graph.add_edge(0, 1, np.ones(3 * 10))
    graph.get_node(0).property.parameters.ContactPatterns.school_holiday_damping = \
        mio.DampingSampling(0.6, 1, 0, 0.0, [0], np.ones(3))
    dn = graph.get_node(0).property.parameters.DynamicNPIsInfectedSymptoms
    ds = mio.DampingSampling(0.4, 0, 0, 0.0, [0], np.ones(3))
    ds.value.set_distribution(mio.ParameterDistributionUniform(0.3, 0.5))
    dn.set_threshold(200.0, [ds])
    return graph


def test_sample_graph():
    """test_parameter_studies.cpp:110-185 and parameter_space.h:211-251: one damping on the edge and on each
    node's contact matrix; node 0's infection parameters, contact dampings and dynamic NPIs go to every node
    and edge; each node keeps its demographics, ICU and test-and-trace capacities and school holidays."""
    mio.seed_random_number_generator(33)
    graph = _sample_graph_setup()
    sampled = osecir.draw_sample(graph)
    assert sampled.num_nodes == 2 and sampled.num_edges == 1
    n0, n1 = (sampled.get_node(v).property for v in range(2))
    p0, p1 = n0.parameters, n1.parameters
    for name in ("TimeExposed", "TimeInfectedSymptoms", "TimeInfectedSevere", "TransmissionProbabilityOnContact",
                 "DeathsPerCritical", "MaxRiskOfInfectionFromSymptomatic"):
        for i in range(3):
            assert getattr(p0, name)[i].value == getattr(p1, name)[i].value, (name, i)
    assert p0.Seasonality.value == p1.Seasonality.value
    # node 1's own capacity and distribution
    assert 30.0 <= p1.ICUCapacity.value <= 50.0
    assert isinstance(p1.ICUCapacity.get_distribution(), mio.ParameterDistributionUniform)
    assert p0.ICUCapacity.value != p1.ICUCapacity.value
    # demographics are per node: node 1's group totals are its own
    assert n1.populations.get_group_total_AgeGroup(0) == pytest.approx(5000.0, abs=1e-6)
    assert n0.populations.get_group_total_AgeGroup(0) == pytest.approx(10000 / 3, abs=1e-6)
    assert n0.populations[0, S.Exposed].value != n1.populations[0, S.Exposed].value
    # contact matrices: node 0 has the one sampled damping; node 1's manual damping is dropped and its
    # school holiday adds the on/off pair of node 0's holiday damping
    d0 = p0.ContactPatterns.cont_freq_mat[0].dampings
    d1 = p1.ContactPatterns.cont_freq_mat[0].dampings
    assert len(d0) == 1 and d0[0].time == 50.0
    assert [d.time for d in d1] == [50.0, 10.0, 20.0]
    np.testing.assert_array_equal(d1[0].coeffs, d0[0].coeffs)
    assert (d1[1].level, d1[1].coeffs[0, 0], d1[2].coeffs[0, 0]) == (1, 0.6, 0.0)
    assert p1.ContactPatterns.school_holidays == [(10.0, 20.0)]
    assert p0.ContactPatterns.school_holidays == []
    # dynamic NPIs of node 0, sampled once, on every node and edge
    v_npi = p0.DynamicNPIsInfectedSymptoms.threshold[0][1][0].value.value
    assert 0.3 <= v_npi <= 0.5 and v_npi != 0.4
    assert p1.DynamicNPIsInfectedSymptoms.threshold[0][1][0].value.value == v_npi
    ep = sampled.get_edge(0).property
    assert ep.dynamic_npis_infected.threshold[0][0] == 200.0
    assert ep.dynamic_npis_infected.threshold[0][1][0].value.value == v_npi
    # the edge: the sampled contact damping as a mobility damping, the damping's value on every compartment
    ed = ep.coefficients[0].dampings
    assert len(ed) == 1 and ed[0].time == 50.0
    v = p0.ContactPatterns.dampings[0].value.value
    np.testing.assert_array_equal(ed[0].coeffs, np.full(30, v))
    # the input graph's edge is left as it was
    assert graph.get_edge(0).property.coefficients[0].dampings == []


def test_graph_sample_without_distributions():
    """With no distributions, draw_sample changes no value but gives every node node 0's parameters (except
    the capacities) and rebuilds the contact matrices, dropping manual dampings."""
    from tests import pymirror_setups as ps
    graph = ps.study_graph(n=3, G=2)
    before = copy.deepcopy(graph)
    sampled = osecir.draw_sample(graph)
    for v in range(3):
        m, b = sampled.get_node(v).property, before.get_node(v).property
        np.testing.assert_array_equal(m.populations.get_compartments(), b.populations.get_compartments())
        b0 = before.get_node(0).property.parameters
        for i in range(2):
            assert m.parameters.TransmissionProbabilityOnContact[i].value == \
                b0.TransmissionProbabilityOnContact[i].value
        np.testing.assert_array_equal(m.parameters.ContactPatterns.cont_freq_mat[0].baseline,
                                      b0.ContactPatterns.cont_freq_mat[0].baseline)
        assert m.parameters.ContactPatterns.cont_freq_mat[0].dampings == []
        assert m.parameters.TestAndTraceCapacity.value == b.parameters.TestAndTraceCapacity.value
    for k in range(sampled.num_edges):
        assert sampled.get_edge(k).property.coefficients[0].dampings == []


def test_sampled_study_graph_is_reproducible():
    graph = _sample_graph_setup()
    mio.seed_random_number_generator(5)
    a = osecir.draw_sample(copy.deepcopy(graph))
    mio.seed_random_number_generator(5)
    b = osecir.draw_sample(copy.deepcopy(graph))
    for v in range(2):
        np.testing.assert_array_equal(a.get_node(v).property.populations.get_compartments(),
                                      b.get_node(v).property.populations.get_compartments())
        assert a.get_node(v).property.parameters.TimeExposed[0].value == \
            b.get_node(v).property.parameters.TimeExposed[0].value


def test_cpp_layer_sampling(tmp_path):
    """The same rules through the C++ mio:: layer (tests/cpp/test_mio_sampling.cpp: distributions,
    UncertainValue, UncertainContactMatrix, set_params_distributions_normal, draw_sample of a model and a
    graph, distributions and sampling inputs in the graph files)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "test_mio_sampling")
    assert os.path.exists(exe), "tests/cpp/test_mio_sampling is built by __graft_entry__.build()"
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # the C++ files hold the distributions; the Python reader takes them as distribution objects
    g = osecir.read_graph(str(tmp_path))
    d = g.get_node(0).property.parameters.TimeExposed[0].get_distribution()
    assert isinstance(d, mio.ParameterDistributionNormal) and d.get_predefined_samples() == [3.3]
    assert isinstance(g.get_node(1).property.populations[2, S.Recovered].get_distribution(),
                      mio.ParameterDistributionLogNormal)
    cp = g.get_node(1).property.parameters.ContactPatterns
    assert cp.school_holidays == [(10.0, 20.0)] and len(cp.dampings) == 1
