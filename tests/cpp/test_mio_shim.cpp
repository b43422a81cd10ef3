// Tests of the mio:: call-shape layer (include/mgo_mio.hpp): code written as against the reference's
// C++ API runs on the device and matches the oracle (oracle/liboracle.so, the CPU restatement of the
// reference algorithm) on the same inputs.
//
// Host-only checks (graph edge order, interpolation) run everywhere. Without a GPU the program prints
// "NO DEVICE" after them and exits 0; with a GPU it prints "ALL PASSED" when every device check holds.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../include/mgo_mio.hpp"
#include "../../oracle/oracle.h"

static int g_fail = 0;

#define CHECK(...)                                                                                               \
    do {                                                                                                         \
        if (!(__VA_ARGS__)) {                                                                                    \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #__VA_ARGS__);                                   \
            ++g_fail;                                                                                            \
        }                                                                                                        \
    } while (0)

using mio::AgeGroup;
using mio::osecir::InfectionState;
using Graph = mio::Graph<mio::SimulationNode<double, mio::osecir::Simulation<double>>, mio::MobilityEdge<double>>;

// max over all entries of |a - b| / max(|b|, 1e-6)
static double rel_err(const double* a, const double* b, size_t n)
{
    double m = 0;
    for (size_t i = 0; i < n; ++i) {
        m = std::max(m, std::abs(a[i] - b[i]) / std::max(std::abs(b[i]), 1e-6));
    }
    return m;
}

// oracle daily values of a graph built by the layer, compared with the nodes' device results
static double compare_graph(const Graph& g, double t0, double dt, double tmax)
{
    mio::detail::GraphArrays a = mio::detail::graph_arrays(g);
    mgo_sim_cfg cfg{};
    cfg.t0         = t0;
    cfg.dt_graph   = dt;
    cfg.dt_node    = a.dt_node;
    cfg.horizon    = tmax;
    cfg.integrator = a.integrator;
    const int cap  = int(std::floor(tmax + 1e-10) - std::ceil(t0 - 1e-10)) + 1;
    const size_t nv = size_t(a.G) * 10;
    std::vector<double> ref(size_t(cap) * size_t(a.n) * nv);
    int32_t nd = 0;
    orc_stats st{};
    if (orc_graph_run(&a.desc, &cfg, tmax, cap, ref.data(), &nd, &st) != 0 || nd != cap) {
        std::printf("oracle failed\n");
        return 1e300;
    }
    double err = 0;
    for (int v = 0; v < a.n; ++v) {
        const auto& ts = g.nodes()[size_t(v)].property.get_result();
        if (int(ts.get_num_time_points()) < nd) {
            std::printf("node %d has %zu time points, expected %d\n", v, ts.get_num_time_points(), nd);
            return 1e300;
        }
        for (int d = 0; d < nd; ++d) {
            CHECK(std::abs(ts.get_time(size_t(d)) - (std::ceil(t0 - 1e-10) + d)) < 1e-12);
            err = std::max(err, rel_err(ts.value_data(size_t(d)), ref.data() + (size_t(d) * size_t(a.n) + size_t(v)) * nv,
                                        nv));
        }
    }
    return err;
}

// the two-node set-up of the reference's graph example (cpp/examples/ode_secir_graph.cpp:30-104)
static Graph example_graph(double t0)
{
    const size_t num_groups = 1;
    mio::osecir::Model<double> model(num_groups);
    model.populations[{AgeGroup(0), InfectionState::Susceptible}] = 10000;
    model.parameters.set<mio::osecir::StartDay<double>>(0);
    model.parameters.set<mio::osecir::Seasonality<double>>(0.2);
    model.parameters.get<mio::osecir::TimeExposed<double>>()            = 3.2;
    model.parameters.get<mio::osecir::TimeInfectedNoSymptoms<double>>() = 2.0;
    model.parameters.get<mio::osecir::TimeInfectedSymptoms<double>>()   = 5.8;
    model.parameters.get<mio::osecir::TimeInfectedSevere<double>>()     = 9.5;
    model.parameters.get<mio::osecir::TimeInfectedCritical<double>>()   = 7.1;
    model.parameters.get<mio::osecir::TransmissionProbabilityOnContact<double>>()  = 0.1;
    model.parameters.get<mio::osecir::RelativeTransmissionNoSymptoms<double>>()    = 0.7;
    model.parameters.get<mio::osecir::RecoveredPerInfectedNoSymptoms<double>>()    = 0.09;
    model.parameters.get<mio::osecir::RiskOfInfectionFromSymptomatic<double>>()    = 0.25;
    model.parameters.get<mio::osecir::MaxRiskOfInfectionFromSymptomatic<double>>() = 0.45;
    model.parameters.get<mio::osecir::TestAndTraceCapacity<double>>()              = 35;
    model.parameters.get<mio::osecir::SeverePerInfectedSymptoms<double>>()         = 0.2;
    model.parameters.get<mio::osecir::CriticalPerSevere<double>>()                 = 0.25;
    model.parameters.get<mio::osecir::DeathsPerCritical<double>>()                 = 0.3;
    mio::ContactMatrixGroup& contact_matrix = model.parameters.get<mio::osecir::ContactPatterns<double>>();
    contact_matrix[0] = mio::ContactMatrix(mio::Matrix::Constant(1, 1, 10));

    auto model_group1 = model;
    auto model_group2 = model;
    mio::ContactMatrixGroup& contact_matrix_m1 = model_group1.parameters.get<mio::osecir::ContactPatterns<double>>();
    contact_matrix_m1[0].add_damping(0.7, mio::SimulationTime(15.0));
    model_group1.populations[{AgeGroup(0), InfectionState::Susceptible}] = 9990;
    model_group1.populations[{AgeGroup(0), InfectionState::Exposed}]     = 100;

    std::vector<std::vector<size_t>> indices_save_edges(2);
    Graph g;
    g.add_node(1001, model_group1, t0);
    g.add_node(1002, model_group2, t0);
    g.add_edge(0, 1, mio::Vector::Constant(size_t(InfectionState::Count), 0.1), indices_save_edges);
    g.add_edge(1, 0, mio::Vector::Constant(size_t(InfectionState::Count), 0.1), indices_save_edges);
    return g;
}

// a random graph with every feature of the device path
static mio::Graph<mio::osecir::Model<double>, mio::MobilityParameters<double>> random_params_graph(int n, int G,
                                                                                                     uint64_t seed)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    mio::Graph<mio::osecir::Model<double>, mio::MobilityParameters<double>> pg;
    for (int v = 0; v < n; ++v) {
        mio::osecir::Model<double> m(G);
        auto& p = m.parameters;
        p.set<mio::osecir::Seasonality<double>>(0.1 + 0.2 * U(rng));
        p.set<mio::osecir::StartDay<double>>(30.0 * U(rng));
        p.get<mio::osecir::TestAndTraceCapacity<double>>() = 20.0 + 100.0 * U(rng);
        p.get<mio::osecir::ICUCapacity<double>>()          = 50.0 + 50.0 * U(rng);
        for (size_t a = 0; a < size_t(G); ++a) {
            p.get<mio::osecir::TimeExposed<double>>()[AgeGroup(a)]            = 2 + 3 * U(rng);
            p.get<mio::osecir::TimeInfectedNoSymptoms<double>>()[AgeGroup(a)] = 2 + 3 * U(rng);
            p.get<mio::osecir::TimeInfectedSymptoms<double>>()[AgeGroup(a)]   = 4 + 4 * U(rng);
            p.get<mio::osecir::TimeInfectedSevere<double>>()[AgeGroup(a)]     = 5 + 5 * U(rng);
            p.get<mio::osecir::TimeInfectedCritical<double>>()[AgeGroup(a)]   = 5 + 5 * U(rng);
            p.get<mio::osecir::TransmissionProbabilityOnContact<double>>()[AgeGroup(a)] = 0.05 + 0.1 * U(rng);
            p.get<mio::osecir::RelativeTransmissionNoSymptoms<double>>()[AgeGroup(a)]   = 0.7;
            p.get<mio::osecir::RecoveredPerInfectedNoSymptoms<double>>()[AgeGroup(a)]   = 0.1 + 0.1 * U(rng);
            p.get<mio::osecir::RiskOfInfectionFromSymptomatic<double>>()[AgeGroup(a)]   = 0.2 + 0.1 * U(rng);
            p.get<mio::osecir::MaxRiskOfInfectionFromSymptomatic<double>>()[AgeGroup(a)] = 0.5;
            p.get<mio::osecir::SeverePerInfectedSymptoms<double>>()[AgeGroup(a)]        = 0.1 + 0.1 * U(rng);
            p.get<mio::osecir::CriticalPerSevere<double>>()[AgeGroup(a)]                = 0.1 + 0.2 * U(rng);
            p.get<mio::osecir::DeathsPerSevere<double>>()[AgeGroup(a)]                  = 0.01 * U(rng);
            p.get<mio::osecir::DeathsPerCritical<double>>()[AgeGroup(a)]                = 0.2 + 0.1 * U(rng);
            const double total = 1e4 + 1e5 * U(rng);
            m.populations[{AgeGroup(a), InfectionState::Exposed}]            = 50 * U(rng);
            m.populations[{AgeGroup(a), InfectionState::InfectedNoSymptoms}] = 20 * U(rng);
            m.populations[{AgeGroup(a), InfectionState::InfectedSymptoms}]   = 10 * U(rng);
            m.populations.set_difference_from_group_total<AgeGroup>({AgeGroup(a), InfectionState::Susceptible},
                                                                    total);
        }
        const size_t Gs = size_t(G);
        mio::Matrix B(Gs, Gs), M(Gs, Gs);
        for (int i = 0; i < G; ++i) {
            for (int j = 0; j < G; ++j) {
                B(size_t(i), size_t(j)) = 1 + 5 * U(rng);
                M(size_t(i), size_t(j)) = 0.2 * U(rng);
            }
        }
        mio::ContactMatrixGroup& cm = p.get<mio::osecir::ContactPatterns<double>>();
        cm[0]    = mio::ContactMatrix(B, M);
        cm[0].add_damping(0.3, mio::DampingLevel(0), mio::DampingType(0), mio::SimulationTime(5.0));
        cm[0].add_damping(0.2, mio::DampingLevel(1), mio::DampingType(0), mio::SimulationTime(9.0));
        cm[0].add_damping(0.5, mio::DampingLevel(0), mio::DampingType(1), mio::SimulationTime(12.0));
        pg.add_node(v, m);
    }
    for (int v = 0; v < n; ++v) {
        for (int k = 0; k < 3; ++k) {
            const int w = int(rng() % uint64_t(n));
            if (w == v) {
                continue;
            }
            mio::Vector c(size_t(10 * G));
            for (size_t i = 0; i < c.size(); ++i) {
                c[i] = 0.02 + 0.05 * U(rng);
            }
            pg.add_edge(size_t(v), size_t(w), c);
        }
    }
    return pg;
}

// ParameterStudy-style sampling: each run scales the contact baselines by a run-specific factor
static Graph sample_graph(const mio::Graph<mio::osecir::Model<double>, mio::MobilityParameters<double>>& pg,
                          double t0, double dt, size_t run_idx, bool fixed)
{
    std::mt19937_64 rng(1000 + run_idx);
    std::uniform_real_distribution<double> U(0.8, 1.2);
    Graph g;
    for (const auto& node : pg.nodes()) {
        auto m   = node.property;
        mio::ContactMatrixGroup& cm = m.parameters.get<mio::osecir::ContactPatterns<double>>();
        auto& B  = cm[0].get_baseline();
        const double f = U(rng);
        for (size_t i = 0; i < B.rows(); ++i) {
            for (size_t j = 0; j < B.cols(); ++j) {
                B(i, j) *= f;
            }
        }
        auto& nd = g.add_node(node.id, m, t0, dt);
        if (fixed) {
            nd.property.get_simulation().set_integrator_core(std::make_unique<mio::ExplicitStepperWrapper<double>>());
        }
    }
    for (const auto& e : pg.edges()) {
        g.add_edge(e.start_node_idx, e.end_node_idx, e.property);
    }
    return g;
}

static void host_checks()
{
    // Graph::add_edge keeps edges sorted by (start, end) and replaces a repeated edge
    mio::Graph<int, double> g;
    for (int i = 0; i < 4; ++i) {
        g.add_node(i, i);
    }
    g.add_edge(2, 1, 1.0);
    g.add_edge(0, 3, 2.0);
    g.add_edge(2, 0, 3.0);
    g.add_edge(0, 3, 4.0);
    CHECK(g.edges().size() == 3);
    CHECK(g.edges()[0].start_node_idx == 0 && g.edges()[0].end_node_idx == 3 && g.edges()[0].property == 4.0);
    CHECK(g.edges()[1].start_node_idx == 2 && g.edges()[1].end_node_idx == 0);
    CHECK(g.edges()[2].start_node_idx == 2 && g.edges()[2].end_node_idx == 1);
    bool threw = false;
    try {
        g.add_edge(0, 7, 1.0);
    }
    catch (const mio::Error& e) {
        threw = e.status == MGO_ERR_OUT_OF_RANGE;
    }
    CHECK(threw);

    // interpolate_simulation_result: linear between points, exact at points within 1e-10
    mio::TimeSeries<double> ts(2);
    const double p0[2] = {0, 10}, p1[2] = {1, 30}, p2[2] = {4, 0};
    ts.add_time_point(-0.5, p0);
    ts.add_time_point(1.0 + 1e-12, p1);
    ts.add_time_point(2.5, p2);
    auto d = mio::interpolate_simulation_result(ts);
    CHECK(d.get_num_time_points() == 3); // days 0, 1, 2
    CHECK(std::abs(d.get_value(0)[0] - (0 + (0.5 / (1.5 + 1e-12)) * 1)) < 1e-12);
    CHECK(d.get_value(1)[1] == 30);
    CHECK(std::abs(d.get_value(2)[0] - (1 + (1.0 - 1e-12) / (1.5 - 1e-12) * 3)) < 1e-12);

    // populations: flat index g*10 + state, set_difference_from_total
    mio::osecir::Model<double> m(2);
    m.populations[{AgeGroup(1), InfectionState::Exposed}] = 5;
    m.populations.set_difference_from_total({AgeGroup(0), InfectionState::Susceptible}, 100);
    CHECK(m.populations.get_flat_index({AgeGroup(1), InfectionState::Exposed}) == 11);
    CHECK(m.populations[{AgeGroup(0), InfectionState::Susceptible}] == 95);
    CHECK(m.parameters.get<mio::osecir::TestAndTraceCapacityMaxRisk<double>>() == 5.0);
}

int main()
{
    host_checks();
    std::printf("host checks: %s\n", g_fail ? "FAILED" : "ok");

    mgo_device_cfg dc{0, 0, 0};
    mgo_context* probe = nullptr;
    const mgo_status st = mgo_context_create(&dc, &probe);
    if (st == MGO_ERR_DEVICE) {
        std::printf("NO DEVICE\n");
        return g_fail ? 1 : 0;
    }
    if (st != MGO_OK) {
        std::printf("FAIL context: %s\n", mgo_last_error(nullptr));
        return 1;
    }
    mgo_context_destroy(probe);

    // 1. the reference example through make_mobility_sim / advance (default Cash-Karp core)
    {
        const double t0 = 0, tmax = 30, dt = 0.5;
        auto sim = mio::make_mobility_sim<double>(t0, dt, example_graph(t0));
        sim.advance(tmax);
        CHECK(sim.get_t() == tmax);
        const double e = compare_graph(sim.get_graph(), t0, dt, tmax);
        std::printf("example graph: max rel err %.3e\n", e);
        CHECK(e < 1e-6);
    }
    // 2. random graph, fixed step, dampings with levels and types, TTC and ICU, advance in pieces
    {
        auto pg  = random_params_graph(20, 3, 7);
        auto sim = mio::make_mobility_sim<double>(0.0, 0.5, sample_graph(pg, 0.0, 0.1, 0, true));
        sim.set_horizon(25.0);
        sim.advance(10.0);
        sim.advance(25.0);
        const double e = compare_graph(sim.get_graph(), 0.0, 0.5, 25.0);
        std::printf("random graph fixed dt: max rel err %.3e\n", e);
        CHECK(e < 1e-12);
    }
    // 3. ParameterStudy::run: one device batch, results per run equal the oracle of the same sample
    {
        auto pg = random_params_graph(12, 2, 11);
        mio::ParameterStudy<decltype(pg), double> study(pg, 0.0, 20.0, 0.5, 16);
        auto results = study.run(
            [](const auto& params, double t0, double dt, size_t run) {
                return mio::make_mobility_sim<double>(t0, dt, sample_graph(params, t0, 0.1, run, false));
            },
            [](auto&& sim, size_t) {
                return std::move(sim).get_graph();
            });
        CHECK(results.size() == 16);
        double e = 0;
        for (auto& g : results) {
            e = std::max(e, compare_graph(g, 0.0, 0.5, 20.0));
        }
        std::printf("parameter study (16 runs): max rel err %.3e\n", e);
        CHECK(e < 1e-6);
        // ensemble statistics of the study on the device equal the reference's arithmetic on the host:
        // run-order sum of value / num_runs, and the sorted run values at floor(num_runs * p)
        auto ens  = mio::interpolate_ensemble_results(results);
        auto mean = mio::ensemble_mean(ens);
        auto p90  = mio::ensemble_percentile(ens, 0.9);
        CHECK(ens.size() == 16 && mean.size() == ens[0].size() && p90.size() == ens[0].size());
        bool same = true;
        for (size_t n = 0; n < mean.size() && same; ++n) {
            for (size_t t = 0; t < mean[n].get_num_time_points(); ++t) {
                for (size_t k = 0; k < mean[n].get_num_elements(); ++k) {
                    double acc = 0.0;
                    std::vector<double> col;
                    for (const auto& run : ens) {
                        acc += run[n].value_data(t)[k] / double(ens.size());
                        col.push_back(run[n].value_data(t)[k]);
                    }
                    std::sort(col.begin(), col.end());
                    same = same && mean[n].value_data(t)[k] == acc &&
                           p90[n].value_data(t)[k] == col[static_cast<size_t>(double(col.size()) * 0.9)];
                }
            }
        }
        std::printf("ensemble mean / 90th percentile on the device: %s\n", same ? "bitwise equal" : "DIFFER");
        CHECK(same);
        auto sums     = mio::sum_nodes(ens);
        bool sums_ok  = sums.size() == ens.size();
        for (size_t r = 0; r < ens.size() && sums_ok; ++r) {
            sums_ok = sums[r].size() == 1 && sums[r][0].get_num_time_points() == ens[r][0].get_num_time_points();
            for (size_t t = 0; t < ens[r][0].get_num_time_points() && sums_ok; ++t) {
                for (size_t k = 0; k < ens[r][0].get_num_elements(); ++k) {
                    double acc = 0.0;
                    for (const auto& node : ens[r]) {
                        acc += node.value_data(t)[k];
                    }
                    sums_ok = sums_ok && sums[r][0].value_data(t)[k] == acc;
                }
            }
        }
        std::printf("sum_nodes on the device: %s\n", sums_ok ? "bitwise equal" : "DIFFER");
        CHECK(sums_ok);
    }
    // 3b. ParameterStudy::run with the reference's sampling in create_simulation (draw_sample of a copy of the
    //     parameter graph, make_sampled_graph_simulation): each run's result equals the oracle of its sample
    {
        auto pg = random_params_graph(10, 2, 13);
        for (auto& node : pg.nodes()) {
            mio::osecir::set_params_distributions_normal(node.property, 0.0, 20.0, 0.2);
        }
        mio::thread_local_rng().seed({7});
        std::vector<double> texp;
        mio::ParameterStudy<decltype(pg), double> study(pg, 0.0, 20.0, 0.5, 8);
        auto results = study.run(
            [&texp](const auto& params, double t0, double dt, size_t) {
                auto copy    = params;
                auto sampled = mio::osecir::draw_sample(copy);
                texp.push_back(sampled.nodes()[0].property.parameters.template get<mio::osecir::TimeExposed<double>>()[AgeGroup(0)]);
                return mio::make_sampled_graph_simulation<double>(sampled, t0, dt, dt);
            },
            [](auto&& sim, size_t) {
                return std::move(sim).get_graph();
            });
        CHECK(results.size() == 8);
        double e = 0;
        for (auto& g : results) {
            e = std::max(e, compare_graph(g, 0.0, 0.5, 20.0));
        }
        std::sort(texp.begin(), texp.end());
        CHECK(std::adjacent_find(texp.begin(), texp.end()) == texp.end()); // every run drew its own sample
        std::printf("sampled parameter study (8 runs): max rel err %.3e\n", e);
        CHECK(e < 1e-6);
    }
    // 4. osecir::simulate with RKIntegratorCore
    {
        auto pg      = random_params_graph(1, 4, 3);
        auto model   = pg.nodes()[0].property;
        auto core    = std::make_unique<mio::RKIntegratorCore<double>>(1e-8, 1e-8, 1e-6, 1.0);
        auto result  = mio::osecir::simulate<double>(0.0, 40.0, 0.3, model, std::move(core));
        mio::detail::GraphArrays a;
        a.add_model(model);
        mio::RKIntegratorCore<double> c2(1e-8, 1e-8, 1e-6, 1.0);
        a.add_core(c2, 0.3);
        a.finish();
        mgo_sim_cfg cfg{};
        cfg.t0 = 0.0, cfg.dt_graph = 0.5, cfg.dt_node = 0.3, cfg.horizon = 40.0, cfg.integrator = a.integrator;
        std::vector<double> ref(41 * 40);
        orc_stats ost{};
        const int nd = orc_simulate_daily(&a.desc, &cfg, 40.0, 41, ref.data(), &ost);
        CHECK(nd == 41);
        CHECK(result.get_num_time_points() == 41);
        double e = 0;
        for (size_t d = 0; d < 41 && d < result.get_num_time_points(); ++d) {
            e = std::max(e, rel_err(result.value_data(d), ref.data() + d * 40, 40));
        }
        std::printf("single model RKF45: max rel err %.3e\n", e);
        CHECK(e < 1e-6);
    }
    // 5. dynamic NPIs (DynamicNPIsInfectedSymptoms) on every node, as the reference's parameter-study examples
    //    set them: thresholds in cases per 100,000, two NPIs on contact matrix 0 with group weights
    {
        auto pg = random_params_graph(16, 3, 21);
        for (auto& node : pg.nodes()) {
            mio::DynamicNPIs<double> npis;
            npis.set_duration(mio::SimulationTime(10.0));
            npis.set_implementation_delay(mio::SimulationTime(2.0));
            npis.set_base_value(100000.0);
            npis.set_threshold(50.0, {mio::DampingSampling<double>(0.4, mio::DampingLevel(0), mio::DampingType(0),
                                                                   mio::SimulationTime(0), {0},
                                                                   mio::Vector::Constant(3, 1.0))});
            npis.set_threshold(300.0, {mio::DampingSampling<double>(0.7, mio::DampingLevel(1), mio::DampingType(0),
                                                                    mio::SimulationTime(0), {0},
                                                                    mio::Vector::Constant(3, 0.8))});
            node.property.parameters.get<mio::osecir::DynamicNPIsInfectedSymptoms<double>>() = npis;
        }
        auto sim = mio::make_mobility_sim<double>(0.0, 0.5, sample_graph(pg, 0.0, 0.1, 3, true));
        sim.advance(30.0);
        const double e = compare_graph(sim.get_graph(), 0.0, 0.5, 30.0);
        std::printf("dynamic NPIs, fixed dt: max rel err %.3e\n", e);
        CHECK(e < 1e-12);
    }
    // 6. damped mobility coefficients (minimum, per-group dampings) and edge dynamic NPIs on every edge, as
    //    osecir::draw_sample gives them (cpp/models/ode_secir/parameter_space.h:240-246)
    {
        auto pg = random_params_graph(14, 2, 31);
        mio::DynamicNPIs<double> npis;
        npis.set_duration(mio::SimulationTime(6.0));
        npis.set_base_value(100000.0);
        npis.set_threshold(40.0, {mio::DampingSampling<double>(0.6, mio::DampingLevel(0), mio::DampingType(0),
                                                               mio::SimulationTime(0), {0},
                                                               mio::Vector::Constant(2, 1.0))});
        for (auto& e : pg.edges()) {
            auto& grp   = e.property.get_coefficients();
            auto& b     = grp[0].get_baseline();
            auto& mn    = grp[0].get_minimum();
            for (size_t i = 0; i < b.size(); ++i) {
                mn[i] = 0.25 * b[i];
            }
            grp.add_damping(0.4, mio::DampingLevel(0), mio::DampingType(0), mio::SimulationTime(4.0));
            grp.add_damping(0.7, mio::DampingLevel(1), mio::DampingType(0), mio::SimulationTime(11.0));
            e.property.set_dynamic_npis_infected(npis);
        }
        auto sim = mio::make_mobility_sim<double>(0.0, 0.5, sample_graph(pg, 0.0, 0.1, 5, true));
        sim.advance(25.0);
        const double e = compare_graph(sim.get_graph(), 0.0, 0.5, 25.0);
        std::printf("mobility dampings and edge NPIs, fixed dt: max rel err %.3e\n", e);
        CHECK(e < 1e-12);
    }
    // 7. errors surface as mio::Error with the C status
    {
        Graph g;
        g.add_node(0, mio::osecir::Model<double>(2), 0.0);
        g.add_node(1, mio::osecir::Model<double>(3), 0.0);
        auto sim  = mio::make_mobility_sim<double>(0.0, 0.5, std::move(g));
        int code  = 0;
        try {
            sim.advance(1.0);
        }
        catch (const mio::Error& e) {
            code = e.status;
        }
        CHECK(code == MGO_ERR_INVALID_VALUE);
        Graph h;
        h.add_node(0, mio::osecir::Model<double>(2), 0.0);
        h.add_node(1, mio::osecir::Model<double>(2), 0.0);
        h.add_edge(0, 1, mio::Vector::Constant(10, 0.1)); // needs 20 values for two age groups
        auto sim2 = mio::make_mobility_sim<double>(0.0, 0.5, std::move(h));
        code      = 0;
        try {
            sim2.advance(1.0);
        }
        catch (const mio::Error& e) {
            code = e.status;
        }
        CHECK(code == MGO_ERR_INVALID_VALUE);
    }
    if (g_fail) {
        std::printf("%d FAILED\n", g_fail);
        return 1;
    }
    std::printf("ALL PASSED\n");
    return 0;
}
