// Parameter sampling through the mio:: layer (include/mgo_mio.hpp), no device needed: distributions,
// UncertainValue, UncertainContactMatrix, set_params_distributions_normal and draw_sample of a model and of a
// parameter graph, after cpp/tests/test_parameter_studies.cpp:34-325 and cpp/tests/test_uncertain.cpp:30-160,
// plus the distribution round trip of the graph files (include/mgo_mio_io.hpp).
// Usage: test_mio_sampling DIR (DIR is a scratch directory for the graph files). Exit status 0 when all
// checks hold.
#include "../../include/mgo_mio_io.hpp"

#include <cmath>
#include <iostream>
#include <string>

namespace
{
int failures = 0;

void check(bool ok, const std::string& what)
{
    if (!ok) {
        std::cerr << "FAILED: " << what << "\n";
        ++failures;
    }
}

constexpr double Q = 2.5758;
using Model        = mio::osecir::Model<double>;
using IS           = mio::osecir::InfectionState;
namespace os       = mio::osecir;

// a distribution whose samples are scripted
class Scripted : public mio::ParameterDistribution
{
public:
    explicit Scripted(std::vector<double> seq)
        : m_seq(std::move(seq))
    {
    }
    double get_rand_sample(mio::RandomNumberGenerator&) override
    {
        const double v = m_seq.front();
        m_seq.erase(m_seq.begin());
        return v;
    }
    std::unique_ptr<mio::ParameterDistribution> clone() const override
    {
        return std::make_unique<Scripted>(*this);
    }

private:
    std::vector<double> m_seq;
};

void test_distributions()
{
    mio::thread_local_rng().seed({11});
    // test_parameter_studies.cpp:187-237
    mio::ParameterDistributionNormal d1;
    d1.set_upper_bound(1);
    d1.set_lower_bound(-1);
    const double demanded = d1.get_standard_dev();
    d1.get_sample(mio::thread_local_rng());
    check(d1.get_standard_dev() < demanded && std::abs(d1.get_standard_dev() - 1.0 / Q) < 1e-15,
          "normal: spread shrinks to fit [lb, ub]");
    mio::ParameterDistributionNormal d2(-1.0, 1.0, 0, d1.get_standard_dev(), Q);
    check(d1.get_lower_bound() == d2.get_lower_bound() && d1.get_upper_bound() == d2.get_upper_bound() &&
              d1.get_mean() == d2.get_mean() && d1.get_standard_dev() == d2.get_standard_dev(),
          "normal: five-argument constructor");
    d2.set_mean(5);
    d2.set_standard_dev(1.5);
    d2.set_lower_bound(1);
    d2.set_upper_bound(10);
    d2.check_quantiles();
    check(d2.get_standard_dev() == 1.5, "normal: a spread that fits is kept");
    double sum = 0.0;
    bool inside = true;
    for (int i = 0; i < 1000; ++i) {
        const double v = d2.get_sample(mio::thread_local_rng());
        inside         = inside && v >= 1.0 && v <= 10.0;
        sum += v;
    }
    check(inside && std::abs(sum / 1000 - 5.0) < 0.15, "normal: samples inside the bounds around the mean");
    mio::ParameterDistributionNormal d3(0.999999999 * 3.0, 1.000000001 * 3.0, 3.0, 0.00000001);
    const double v3 = d3.get_sample(mio::thread_local_rng());
    check(std::abs(v3 - 3.0) < 1e-7 && (v3 == d3.get_lower_bound() || v3 == d3.get_upper_bound()),
          "normal: narrow interval clamps after ten redraws");
    mio::ParameterDistributionNormal mean_out(2.0, 10.0, 20.0, 1.0, Q);
    check(mean_out.get_mean() == 4.0, "normal: a mean outside the bounds becomes (ub - lb) / 2");
    mio::ParameterDistributionNormal four(1.0, 5.0, 2.0, 2.0);
    check(std::abs(four.get_standard_dev() - 0.5) < 1e-15, "normal: (lb, ub, mean, quantile)");

    // test_parameter_studies.cpp:239-278
    mio::ParameterDistributionUniform u(1.0, 10.0);
    inside = true;
    for (int i = 0; i < 1000; ++i) {
        const double v = u.get_sample(mio::thread_local_rng());
        inside         = inside && v >= 1.0 && v <= 10.0;
    }
    check(inside && u.get_lower_bound() == 1.0 && u.get_upper_bound() == 10.0, "uniform");
    check(mio::ParameterDistributionConstant(2.0).get_constant() == 2.0, "constant");
    check(mio::ParameterDistributionLogNormal(0.0, 0.25).get_log_stddev() == 0.25, "lognormal");
    check(mio::ParameterDistributionExponential(1.0).get_rate() == 1.0, "exponential");

    // test_parameter_studies.cpp:280-325
    u.add_predefined_sample(2);
    check(u.get_sample(mio::thread_local_rng()) == 2, "predefined sample first");
    check(u.get_sample(mio::thread_local_rng()) != 2, "predefined sample used once");
    mio::ParameterDistributionConstant c(3.0);
    c.add_predefined_sample(1.0);
    check(c.get_sample(mio::thread_local_rng()) == 1.0 && c.get_sample(mio::thread_local_rng()) == 3.0,
          "constant with a predefined sample");

    // the seeded generator repeats
    mio::ParameterDistributionNormal n(0.0, 10.0, 5.0, 1.0, Q);
    mio::thread_local_rng().seed({1, 2, 3});
    const double a = n.get_sample(mio::thread_local_rng());
    mio::thread_local_rng().seed({1, 2, 3});
    check(n.get_sample(mio::thread_local_rng()) == a, "seeded generator repeats");
}

void test_uncertain_value()
{
    // test_uncertain.cpp:30-114
    mio::UncertainValue<double> val(3.0);
    check(val == 3.0, "uv value");
    val = 2.0;
    check(val == 2.0, "uv assign");
    val.set_distribution(mio::ParameterDistributionNormal(0.96, 3.04, 2.0, 0.4));
    mio::UncertainValue<double> val2(val);
    check(val2 == 2.0 && val2.get_distribution().get() != val.get_distribution().get() &&
              val2.get_distribution().get() != nullptr,
          "uv copy has its own distribution");
    mio::UncertainValue<double> sc(2.0);
    sc.set_distribution(Scripted({0.5, 1.5}));
    sc.draw_sample();
    check(sc == 0.5, "uv draw 1");
    sc = 1.0;
    check(sc == 1.0 && sc.get_distribution(), "uv assign keeps the distribution");
    sc.draw_sample();
    check(sc == 1.5, "uv draw 2");
    double& dval = val;
    dval         = 4.0;
    check(val == 4.0, "uv as double&");
    mio::UncertainValue<double> val3;
    val3 = val;
    check(val3 == val && val3.get_distribution(), "uv assignment copies the distribution");
    mio::UncertainValue<double> p(3.0);
    p.set_distribution(mio::ParameterDistributionNormal(1.44, 4.56, 3.0, 0.6));
    p.get_distribution()->add_predefined_sample(17111.0);
    mio::UncertainValue<double> p2(p);
    p.draw_sample();
    check(p == 17111.0, "predefined 1");
    p.draw_sample();
    check(p != 17111.0, "predefined 2");
    p2.draw_sample();
    check(p2 == 17111.0, "predefined copied");
    mio::UncertainValue<double> none(4.0);
    check(none.draw_sample() == 4.0, "no distribution: value kept");
}

void test_uncertain_matrix()
{
    // test_uncertain.cpp:116-160
    mio::Matrix B(2, 2);
    for (size_t i = 0; i < 2; ++i) {
        for (size_t j = 0; j < 2; ++j) {
            B(i, j) = double((i + 1) * (j + 1));
        }
    }
    mio::ContactMatrix cm(B);
    cm.add_damping(0.7, mio::SimulationTime(30.));
    mio::ContactMatrixGroup grp(1, 2);
    grp[0] = cm;
    mio::UncertainContactMatrix<double> um(grp);
    check(um.get_cont_freq_mat()[0].get_baseline()(0, 1) == 2 && um.get_cont_freq_mat()[0].get_dampings().size() == 1,
          "uncertain matrix baseline");
    um.get_dampings().emplace_back(mio::UncertainValue<double>(0.5), mio::DampingLevel(0), mio::DampingType(0),
                                   mio::SimulationTime(3.0), std::vector<size_t>(1, 0), mio::Vector(2, 1.0));
    mio::Vector hw(2, 0.0);
    hw[0]                            = 1.0;
    um.get_school_holiday_damping() = mio::DampingSampling<double>(
        mio::UncertainValue<double>(1.), mio::DampingLevel(1), mio::DampingType(0), mio::SimulationTime(0.0),
        std::vector<size_t>(1, 0), hw);
    um.get_school_holidays().assign({{mio::SimulationTime(5.0), mio::SimulationTime(17.0)}});
    um.draw_sample(true);
    check(um.get_cont_freq_mat()[0].get_dampings().size() == 4, "accumulated dampings");
    um.draw_sample();
    const auto& ds = um.get_cont_freq_mat()[0].get_dampings();
    check(ds.size() == 3, "replaced dampings");
    check(ds[0].level == 0 && ds[0].coeffs(0, 0) == 0.5 && ds[0].time == 3.0, "sampled damping");
    check(ds[1].level == 1 && ds[1].coeffs(0, 0) == 1.0 && ds[1].coeffs(1, 1) == 0.0 && ds[1].time == 5.0,
          "holiday on");
    check(ds[2].level == 1 && ds[2].coeffs(0, 0) == 0.0 && ds[2].time == 17.0, "holiday off");
    check(double(um.get_school_holiday_damping().get_value()) == 1.0, "holiday damping keeps its value");
}

// the model of test_parameter_studies.cpp:34-95
Model secir_model(size_t G = 3)
{
    Model model{int(G)};
    const double fact = 1.0 / double(G);
    auto& params      = model.parameters;
    for (size_t g = 0; g < G; ++g) {
        const mio::AgeGroup i(g);
        params.get<os::TimeExposed<double>>()[i]            = 2.6;
        params.get<os::TimeInfectedNoSymptoms<double>>()[i] = 2.6;
        params.get<os::TimeInfectedSymptoms<double>>()[i]   = 5.;
        params.get<os::TimeInfectedSevere<double>>()[i]     = 10.;
        params.get<os::TimeInfectedCritical<double>>()[i]   = 8.;
        params.get<os::Seasonality<double>>()               = 0.0;
        params.get<os::ICUCapacity<double>>()               = 100.0;
        params.get<os::TestAndTraceCapacity<double>>()      = 10.0;
        model.populations[{i, IS::Exposed}]                 = fact * 100;
        model.populations[{i, IS::InfectedNoSymptoms}]      = fact * 50;
        model.populations[{i, IS::InfectedSymptoms}]        = fact * 50;
        model.populations[{i, IS::InfectedSevere}]          = fact * 20;
        model.populations[{i, IS::InfectedCritical}]        = fact * 10;
        model.populations[{i, IS::Recovered}]               = fact * 10;
        model.populations[{i, IS::Dead}]                    = 0;
        model.populations.set_difference_from_group_total<mio::AgeGroup>({i, IS::Susceptible}, fact * 10000);
        params.get<os::TransmissionProbabilityOnContact<double>>()[i]  = 0.05;
        params.get<os::RelativeTransmissionNoSymptoms<double>>()[i]    = 0.67;
        params.get<os::RecoveredPerInfectedNoSymptoms<double>>()[i]    = 0.09;
        params.get<os::RiskOfInfectionFromSymptomatic<double>>()[i]    = 0.25;
        params.get<os::MaxRiskOfInfectionFromSymptomatic<double>>()[i] = 0.85;
        params.get<os::SeverePerInfectedSymptoms<double>>()[i]         = 0.2;
        params.get<os::CriticalPerSevere<double>>()[i]                 = 0.25;
        params.get<os::DeathsPerCritical<double>>()[i]                 = 0.3;
    }
    mio::ContactMatrixGroup& contact_matrix = params.get<os::ContactPatterns<double>>();
    contact_matrix[0] = mio::ContactMatrix(mio::Matrix::Constant(G, G, fact * 10));
    return model;
}

bool in_bounds(const mio::UncertainValue<double>& v)
{
    auto d = v.get_distribution();
    if (auto* n = dynamic_cast<const mio::ParameterDistributionNormal*>(d.get())) {
        return v >= n->get_lower_bound() && v <= n->get_upper_bound();
    }
    return true;
}

void test_sample_from_secir_params()
{
    // test_parameter_studies.cpp:34-108
    mio::thread_local_rng().seed({21});
    Model model = secir_model();
    os::set_params_distributions_normal(model, 0, 100, 0.2);
    auto& params = model.parameters;
    {
        auto* d = dynamic_cast<mio::ParameterDistributionNormal*>(
            params.get<os::TimeInfectedSevere<double>>()[mio::AgeGroup(1)].get_distribution().get());
        check(d && std::abs(d->get_lower_bound() - 4.8) < 1e-12 && std::abs(d->get_upper_bound() - 15.2) < 1e-12 &&
                  d->get_mean() == 10.0 && std::abs(d->get_quantile() - 2.0) < 1e-15,
              "set_params_distributions_normal bounds");
        check(dynamic_cast<mio::ParameterDistributionConstant*>(
                  params.get<os::Seasonality<double>>().get_distribution().get()) != nullptr,
              "seasonality 0: constant distribution");
        check(!model.populations[{mio::AgeGroup(0), IS::Susceptible}].get_distribution() &&
                  !model.populations[{mio::AgeGroup(0), IS::Dead}].get_distribution(),
              "S and D are not sampled");
        check(params.get<os::ContactPatterns<double>>().get_dampings().size() == 1 &&
                  params.get<os::ContactPatterns<double>>().get_dampings()[0].get_time().get() == 50.0,
              "one contact damping sampling at mid-time");
    }
    params.get<os::ContactPatterns<double>>().get_cont_freq_mat().add_damping(0.3, mio::SimulationTime(5.0));
    const double te_before = params.get<os::TimeExposed<double>>()[mio::AgeGroup(0)];
    os::draw_sample(model);
    for (size_t g = 0; g < 3; ++g) {
        const mio::AgeGroup i(g);
        check(params.get<os::RelativeTransmissionNoSymptoms<double>>()[mio::AgeGroup(0)] ==
                      params.get<os::RelativeTransmissionNoSymptoms<double>>()[i] &&
                  params.get<os::RiskOfInfectionFromSymptomatic<double>>()[mio::AgeGroup(0)] ==
                      params.get<os::RiskOfInfectionFromSymptomatic<double>>()[i] &&
                  params.get<os::MaxRiskOfInfectionFromSymptomatic<double>>()[mio::AgeGroup(0)] ==
                      params.get<os::MaxRiskOfInfectionFromSymptomatic<double>>()[i] &&
                  params.get<os::TimeExposed<double>>()[mio::AgeGroup(0)] == params.get<os::TimeExposed<double>>()[i],
              "shared parameters equal across groups");
        check(std::abs(model.populations.get_group_total(i) - 10000.0 / 3) < 1e-6, "group totals kept");
        check(params.get<os::TransmissionProbabilityOnContact<double>>()[i] >= 0, "probability >= 0");
        check(in_bounds(params.get<os::TimeInfectedSevere<double>>()[i]) &&
                  in_bounds(model.populations[{i, IS::Exposed}]),
              "samples within their bounds");
    }
    check(params.get<os::TimeExposed<double>>()[mio::AgeGroup(0)] != te_before, "TimeExposed was drawn");
    const auto& cds = params.get<os::ContactPatterns<double>>().get_cont_freq_mat()[0].get_dampings();
    check(cds.size() == 1 && cds[0].time == 50.0, "the manual damping is replaced by the sampled one");
}

void test_apply_constraints()
{
    Model m = secir_model();
    m.parameters.get<os::CriticalPerSevere<double>>()[mio::AgeGroup(1)].set_distribution(
        mio::ParameterDistributionUniform(1.5, 2.0));
    m.parameters.get<os::DeathsPerSevere<double>>()[mio::AgeGroup(2)] = 0.8;
    os::draw_sample(m);
    check(m.parameters.get<os::CriticalPerSevere<double>>()[mio::AgeGroup(1)] == 0.0, "probability above 1 -> 0");
    check(m.parameters.get<os::DeathsPerSevere<double>>()[mio::AgeGroup(2)] == 0.0,
          "CriticalPerSevere + DeathsPerSevere > 1 -> DeathsPerSevere = 0");
    check(m.parameters.get<os::CriticalPerSevere<double>>()[mio::AgeGroup(0)] == 0.25, "others untouched");
}

using ParamGraph = mio::Graph<Model, mio::MobilityParameters<double>>;

ParamGraph sample_graph_setup()
{
    Model model = secir_model();
    os::set_params_distributions_normal(model, 0, 100, 0.2);
    ParamGraph graph;
    Model m1 = model;
    m1.parameters.get<os::ICUCapacity<double>>() = 40.0;
    m1.parameters.get<os::ICUCapacity<double>>().set_distribution(mio::ParameterDistributionUniform(30.0, 50.0));
    m1.parameters.get<os::ContactPatterns<double>>().get_school_holidays().assign(
        {{mio::SimulationTime(10.0), mio::SimulationTime(20.0)}});
    m1.parameters.get<os::ContactPatterns<double>>().get_cont_freq_mat().add_damping(0.9, mio::SimulationTime(1.0));
    model.parameters.get<os::ContactPatterns<double>>().get_school_holiday_damping() = mio::DampingSampling<double>(
        mio::UncertainValue<double>(0.6), mio::DampingLevel(1), mio::DampingType(0), mio::SimulationTime(0.0),
        std::vector<size_t>(1, 0), mio::Vector(3, 1.0));
    mio::DampingSampling<double> npi(mio::UncertainValue<double>(0.4), mio::DampingLevel(0), mio::DampingType(0),
                                     mio::SimulationTime(0.0), std::vector<size_t>(1, 0), mio::Vector(3, 1.0));
    npi.get_value().set_distribution(mio::ParameterDistributionUniform(0.3, 0.5));
    model.parameters.get<os::DynamicNPIsInfectedSymptoms<double>>().set_threshold(200.0, {npi});
    graph.add_node(0, model);
    graph.add_node(1, m1);
    graph.add_edge(0, 1, mio::MobilityParameters<double>(mio::Vector(3 * 10, 1.0)));
    // every node needs out-edges for the graph files (the reference reads an edge file per node)
    graph.add_edge(1, 0, mio::MobilityParameters<double>(mio::Vector(3 * 10, 0.5)));
    return graph;
}

void test_sample_graph()
{
    // test_parameter_studies.cpp:110-185 and parameter_space.h:211-251
    mio::thread_local_rng().seed({33});
    ParamGraph graph   = sample_graph_setup();
    ParamGraph sampled = os::draw_sample(graph);
    check(sampled.nodes().size() == 2 && sampled.edges().size() == 2, "sampled graph shape");
    const auto& p0 = sampled.nodes()[0].property.parameters;
    const auto& p1 = sampled.nodes()[1].property.parameters;
    bool shared    = true;
    for (size_t g = 0; g < 3; ++g) {
        const mio::AgeGroup i(g);
        shared = shared && p0.get<os::TimeInfectedSevere<double>>()[i] == p1.get<os::TimeInfectedSevere<double>>()[i] &&
                 p0.get<os::TransmissionProbabilityOnContact<double>>()[i] ==
                     p1.get<os::TransmissionProbabilityOnContact<double>>()[i];
    }
    check(shared && p0.get<os::Seasonality<double>>() == p1.get<os::Seasonality<double>>(),
          "node 0's infection parameters on every node");
    const double icu1 = p1.get<os::ICUCapacity<double>>();
    check(icu1 >= 30.0 && icu1 <= 50.0 && icu1 != double(p0.get<os::ICUCapacity<double>>()),
          "each node keeps its own capacity");
    const auto& d0 = p0.get<os::ContactPatterns<double>>().get_cont_freq_mat()[0].get_dampings();
    const auto& d1 = p1.get<os::ContactPatterns<double>>().get_cont_freq_mat()[0].get_dampings();
    check(d0.size() == 1 && d0[0].time == 50.0, "node 0: one sampled contact damping");
    check(d1.size() == 3 && d1[0].time == 50.0 && d1[1].time == 10.0 && d1[2].time == 20.0 &&
              d1[1].coeffs(0, 0) == 0.6 && d1[2].coeffs(0, 0) == 0.0 && d1[0].coeffs == d0[0].coeffs,
          "node 1: sampled damping, then its holidays with node 0's holiday damping");
    const auto& n0  = p0.get<os::DynamicNPIsInfectedSymptoms<double>>().get_thresholds();
    const auto& n1  = p1.get<os::DynamicNPIsInfectedSymptoms<double>>().get_thresholds();
    const double vn = n0.at(0).second.at(0).get_value();
    check(vn >= 0.3 && vn <= 0.5 && vn != 0.4 && double(n1.at(0).second.at(0).get_value()) == vn,
          "dynamic NPIs sampled once, on every node");
    const auto& ep = sampled.edges()[0].property;
    const auto& ed = ep.get_coefficients()[0].get_dampings();
    const double v = p0.get<os::ContactPatterns<double>>().get_dampings()[0].get_value();
    check(ed.size() == 1 && ed[0].time == 50.0 && ed[0].coeffs.size() == 30 && ed[0].coeffs[0] == v &&
              ed[0].coeffs[29] == v,
          "edge: the sampled contact damping as a mobility damping");
    check(double(ep.get_dynamic_npis_infected().get_thresholds().at(0).second.at(0).get_value()) == vn,
          "edge: node 0's dynamic NPIs");
    check(graph.edges()[0].property.get_coefficients()[0].get_dampings().empty(), "input edge unchanged");
    check(std::abs(sampled.nodes()[1].property.populations.get_group_total(mio::AgeGroup(0)) - 10000.0 / 3) < 1e-6,
          "node totals kept");
}

void test_distribution_files(const std::string& dir)
{
    // graph files keep distributions and the contact patterns' sampling inputs
    ParamGraph graph = sample_graph_setup();
    graph.nodes()[0].property.parameters.get<os::TimeExposed<double>>()[mio::AgeGroup(0)]
        .get_distribution()
        ->add_predefined_sample(3.3);
    graph.nodes()[1].property.populations[{mio::AgeGroup(2), IS::Recovered}].set_distribution(
        mio::ParameterDistributionLogNormal(0.5, 0.1));
    auto st = mio::write_graph(graph, dir);
    check(bool(st), "write_graph: " + st.error().formatted_message());
    auto r = mio::read_graph<double, Model>(dir);
    check(bool(r), "read_graph: " + (r ? std::string() : r.error().formatted_message()));
    if (!r) {
        return;
    }
    const auto& g = r.value();
    const auto& te =
        g.nodes()[0].property.parameters.get<os::TimeExposed<double>>()[mio::AgeGroup(0)].get_distribution();
    auto* tn = dynamic_cast<const mio::ParameterDistributionNormal*>(te.get());
    auto* to = dynamic_cast<const mio::ParameterDistributionNormal*>(
        graph.nodes()[0].property.parameters.get<os::TimeExposed<double>>()[mio::AgeGroup(0)].get_distribution().get());
    check(tn && to && tn->get_lower_bound() == to->get_lower_bound() && tn->get_upper_bound() == to->get_upper_bound() &&
              tn->get_mean() == to->get_mean() && tn->get_quantile() == to->get_quantile() &&
              std::abs(tn->get_standard_dev() - to->get_standard_dev()) <= 1e-15 * to->get_standard_dev() &&
              tn->get_predefined_samples() == std::vector<double>{3.3},
          "normal distribution round trip");
    check(dynamic_cast<const mio::ParameterDistributionLogNormal*>(
              g.nodes()[1].property.populations[{mio::AgeGroup(2), IS::Recovered}].get_distribution().get()) != nullptr,
          "lognormal round trip");
    const auto& ucm = g.nodes()[1].property.parameters.get<os::ContactPatterns<double>>();
    check(ucm.get_school_holidays().size() == 1 && ucm.get_school_holidays()[0].second.get() == 20.0 &&
              ucm.get_dampings().size() == 1 && ucm.get_dampings()[0].get_value().get_distribution(),
          "contact sampling inputs round trip");
    check(double(g.nodes()[0].property.parameters.get<os::ContactPatterns<double>>()
                     .get_school_holiday_damping()
                     .get_value()) == 0.6,
          "holiday damping round trip");
    // sampling the graph read back gives the same kind of sample
    mio::thread_local_rng().seed({5});
    auto a = os::draw_sample(graph);
    mio::thread_local_rng().seed({5});
    auto rg = r.value();
    auto b  = os::draw_sample(rg);
    // the predefined sample of the written graph was added after sample_graph_setup; both have it
    check(double(a.nodes()[0].property.parameters.get<os::TimeExposed<double>>()[mio::AgeGroup(0)]) == 3.3 &&
              double(b.nodes()[0].property.parameters.get<os::TimeExposed<double>>()[mio::AgeGroup(0)]) == 3.3,
          "predefined samples survive the files");
    check(a.nodes()[1].property.populations.get_compartments() == b.nodes()[1].property.populations.get_compartments(),
          "same seed, same file: same sample");
}

// ensemble_mean / ensemble_percentile refuse non-uniform [run][node] input before anything reaches the device
void test_ensemble_input_checks()
{
    auto series = [](size_t points) {
        mio::TimeSeries<double> ts(2);
        const double v[2] = {1.0, 2.0};
        for (size_t t = 0; t < points; ++t) {
            ts.add_time_point(double(t), v);
        }
        return ts;
    };
    auto refused = [](auto&& f) {
        try {
            f();
        }
        catch (const mio::Error& e) {
            return e.status == MGO_ERR_INVALID_VALUE;
        }
        return false;
    };
    std::vector<std::vector<mio::TimeSeries<double>>> nodes_differ{{series(3), series(3)}, {series(3)}};
    std::vector<std::vector<mio::TimeSeries<double>>> points_differ{{series(3)}, {series(4)}};
    check(refused([&] { mio::ensemble_mean(nodes_differ); }), "ensemble_mean: runs with different node counts");
    check(refused([&] { mio::ensemble_percentile(points_differ, 0.5); }),
          "ensemble_percentile: different time points");
    check(refused([&] { mio::ensemble_mean(std::vector<std::vector<mio::TimeSeries<double>>>{}); }),
          "ensemble_mean: empty");
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::cerr << "usage: test_mio_sampling DIR\n";
        return 2;
    }
    test_distributions();
    test_uncertain_value();
    test_uncertain_matrix();
    test_sample_from_secir_params();
    test_apply_constraints();
    test_sample_graph();
    test_distribution_files(argv[1]);
    test_ensemble_input_checks();
    if (failures) {
        std::cerr << failures << " check(s) failed\n";
        return 1;
    }
    std::cout << "all sampling checks passed\n";
    return 0;
}
