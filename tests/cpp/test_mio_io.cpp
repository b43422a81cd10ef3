// Parameter-graph JSON files through the mio:: layer (include/mgo_mio_io.hpp), no device needed.
//
//   test_mio_io example DIR      builds the graph of the reference's ode_secir_read_graph.cpp example (plus
//                                dampings, a dynamic NPI and a damped edge), writes it, reads it back and
//                                checks every field; the files stay in DIR for the Python reader
//   test_mio_io roundtrip IN OUT reads the graph in IN (e.g. written by the Python mirror) and writes it to OUT
//   test_mio_io errors DIR       error statuses of read_graph (missing directory, edge file, malformed files)
#include "../../include/mgo_mio_io.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>

namespace
{
int failures = 0;

void check(bool ok, const std::string& what)
{
    if (!ok) {
        std::cout << "FAIL: " << what << "\n";
        ++failures;
    }
}

using Graph = mio::Graph<mio::osecir::Model<double>, mio::MobilityParameters<double>>;

Graph example_graph()
{
    // cpp/examples/ode_secir_read_graph.cpp:62-130 with a 3-node stand-in for the commuter matrix
    const double cont_freq = 10, nb_total_t0 = 10000, nb_exp_t0 = 100, nb_inf_t0 = 50, nb_car_t0 = 50,
                 nb_hosp_t0 = 20, nb_icu_t0 = 10, nb_rec_t0 = 10, nb_dead_t0 = 0;
    mio::osecir::Model<double> model(2);
    const mio::AgeGroup nb_groups = model.parameters.get_num_groups();
    const double fact             = 1.0 / double(nb_groups.get());
    auto& params                  = model.parameters;
    params.set<mio::osecir::ICUCapacity<double>>(std::numeric_limits<double>::max());
    params.set<mio::osecir::StartDay<double>>(0);
    params.set<mio::osecir::Seasonality<double>>(0.1);
    using S = mio::osecir::InfectionState;
    for (auto i = mio::AgeGroup(0); i < nb_groups; ++i) {
        params.get<mio::osecir::TimeExposed<double>>()[i]            = 3.2;
        params.get<mio::osecir::TimeInfectedNoSymptoms<double>>()[i] = 2.0;
        params.get<mio::osecir::TimeInfectedSymptoms<double>>()[i]   = 6.;
        params.get<mio::osecir::TimeInfectedSevere<double>>()[i]     = 12;
        params.get<mio::osecir::TimeInfectedCritical<double>>()[i]   = 8;
        model.populations[{i, S::Exposed}]                            = fact * nb_exp_t0;
        model.populations[{i, S::InfectedNoSymptoms}]                 = fact * nb_car_t0;
        model.populations[{i, S::InfectedSymptoms}]                   = fact * nb_inf_t0;
        model.populations[{i, S::InfectedSevere}]                     = fact * nb_hosp_t0;
        model.populations[{i, S::InfectedCritical}]                   = fact * nb_icu_t0;
        model.populations[{i, S::Recovered}]                          = fact * nb_rec_t0;
        model.populations[{i, S::Dead}]                               = fact * nb_dead_t0;
        model.populations.set_difference_from_group_total<mio::AgeGroup>({i, S::Susceptible}, fact * nb_total_t0);
        params.get<mio::osecir::TransmissionProbabilityOnContact<double>>()[i] = 0.05;
        params.get<mio::osecir::RelativeTransmissionNoSymptoms<double>>()[i]   = 0.67;
        params.get<mio::osecir::RecoveredPerInfectedNoSymptoms<double>>()[i]   = 0.09;
        params.get<mio::osecir::RiskOfInfectionFromSymptomatic<double>>()[i]   = 0.25;
        params.get<mio::osecir::SeverePerInfectedSymptoms<double>>()[i]        = 0.2;
        params.get<mio::osecir::CriticalPerSevere<double>>()[i]                = 0.25;
        params.get<mio::osecir::DeathsPerCritical<double>>()[i]                = 0.3;
    }
    mio::ContactMatrixGroup& contact_matrix = params.get<mio::osecir::ContactPatterns<double>>();
    contact_matrix[0] = mio::ContactMatrix(mio::Matrix::Constant(2, 2, fact * cont_freq));
    contact_matrix[0].get_minimum() = mio::Matrix::Constant(2, 2, 0.1);
    contact_matrix.add_damping(0.3, mio::DampingLevel(1), mio::DampingType(2), mio::SimulationTime(12.5));
    auto& npis = params.get<mio::osecir::DynamicNPIsInfectedSymptoms<double>>();
    npis.set_duration(mio::SimulationTime(10.0));
    npis.set_implementation_delay(mio::SimulationTime(1.5));
    npis.set_base_value(1e5);
    npis.set_threshold(200.0, {mio::DampingSampling<double>(0.4, mio::DampingLevel(0), mio::DampingType(0),
                                                            mio::SimulationTime(0.0), {0},
                                                            mio::Vector::Constant(2, 1.0))});

    const double commuters[3][3] = {{0, 120, 30}, {80, 0, 0}, {10, 40, 0}};
    Graph graph;
    for (int node = 0; node < 3; ++node) {
        graph.add_node(node, model);
    }
    for (int row = 0; row < 3; ++row) {
        for (int col = 0; col < 3; ++col) {
            graph.add_edge(row, col,
                           mio::Vector::Constant(10 * nb_groups.get(),
                                                 commuters[row][col] / graph.nodes()[row].property.populations.get_total()));
        }
    }
    auto& ep = graph.edges()[1].property;
    ep.get_coefficients()[0].get_minimum() = mio::Vector::Constant(20, 0.001);
    ep.get_coefficients().add_damping(0.5, mio::DampingLevel(0), mio::DampingType(0), mio::SimulationTime(4.0));
    return graph;
}

bool same_npis(const mio::DynamicNPIs<double>& a, const mio::DynamicNPIs<double>& b)
{
    if (a.get_thresholds().size() != b.get_thresholds().size() || a.get_duration().get() != b.get_duration().get() ||
        a.get_implementation_delay().get() != b.get_implementation_delay().get() ||
        a.get_base_value() != b.get_base_value() || a.get_directive_begin().get() != b.get_directive_begin().get() ||
        a.get_directive_end().get() != b.get_directive_end().get()) {
        return false;
    }
    for (size_t k = 0; k < a.get_thresholds().size(); ++k) {
        const auto &x = a.get_thresholds()[k], &y = b.get_thresholds()[k];
        if (x.first != y.first || x.second.size() != y.second.size()) {
            return false;
        }
        for (size_t d = 0; d < x.second.size(); ++d) {
            const auto &p = x.second[d], &q = y.second[d];
            if (p.get_value() != q.get_value() || p.get_level().get() != q.get_level().get() ||
                p.get_type().get() != q.get_type().get() || p.get_time().get() != q.get_time().get() ||
                p.get_matrix_indices() != q.get_matrix_indices() ||
                p.get_group_weights().size() != q.get_group_weights().size()) {
                return false;
            }
            for (size_t g = 0; g < p.get_group_weights().size(); ++g) {
                if (p.get_group_weights()[g] != q.get_group_weights()[g]) {
                    return false;
                }
            }
        }
    }
    return true;
}

bool same_vector(const mio::Vector& a, const mio::Vector& b)
{
    if (a.size() != b.size()) {
        return false;
    }
    for (size_t i = 0; i < a.size(); ++i) {
        if (a[i] != b[i]) {
            return false;
        }
    }
    return true;
}

void compare(const Graph& a, const Graph& b)
{
    check(a.nodes().size() == b.nodes().size(), "node count");
    check(a.edges().size() == b.edges().size(), "edge count");
    for (size_t v = 0; v < std::min(a.nodes().size(), b.nodes().size()); ++v) {
        const auto &ma = a.nodes()[v].property, &mb = b.nodes()[v].property;
        check(a.nodes()[v].id == b.nodes()[v].id, "node id");
        check(ma.parameters.scalars()[MGO_P_START_DAY] == mb.parameters.scalars()[MGO_P_START_DAY] &&
                  ma.parameters.scalars()[MGO_P_SEASONALITY] == mb.parameters.scalars()[MGO_P_SEASONALITY] &&
                  ma.parameters.scalars()[MGO_P_ICU_CAPACITY] == mb.parameters.scalars()[MGO_P_ICU_CAPACITY] &&
                  ma.parameters.scalars()[MGO_P_TEST_AND_TRACE_CAPACITY] ==
                      mb.parameters.scalars()[MGO_P_TEST_AND_TRACE_CAPACITY] &&
                  ma.parameters.scalars()[MGO_P_TEST_AND_TRACE_MAX_RISK] ==
                      mb.parameters.scalars()[MGO_P_TEST_AND_TRACE_MAX_RISK],
              "scalar parameters");
        for (int k = 0; k < MGO_NUM_AGE_PARAMS; ++k) {
            for (size_t g = 0; g < ma.get_num_groups(); ++g) {
                check(ma.parameters.age_param(k)[mio::AgeGroup(g)] == mb.parameters.age_param(k)[mio::AgeGroup(g)],
                      "age parameter " + std::to_string(k));
            }
        }
        check(ma.populations.get_compartments() == mb.populations.get_compartments(), "populations");
        const auto& ca = ma.parameters.get<mio::osecir::ContactPatterns<double>>().get_cont_freq_mat();
        const auto& cb = mb.parameters.get<mio::osecir::ContactPatterns<double>>().get_cont_freq_mat();
        check(ca.get_num_matrices() == cb.get_num_matrices(), "contact matrices");
        for (size_t k = 0; k < ca.get_num_matrices(); ++k) {
            check(ca[k].get_baseline() == cb[k].get_baseline() && ca[k].get_minimum() == cb[k].get_minimum(),
                  "contact baseline / minimum");
            check(ca[k].get_dampings().size() == cb[k].get_dampings().size(), "contact dampings");
            for (size_t d = 0; d < std::min(ca[k].get_dampings().size(), cb[k].get_dampings().size()); ++d) {
                const auto &x = ca[k].get_dampings()[d], &y = cb[k].get_dampings()[d];
                check(x.coeffs == y.coeffs && x.level == y.level && x.type == y.type && x.time == y.time,
                      "contact damping");
            }
        }
        check(same_npis(ma.parameters.get<mio::osecir::DynamicNPIsInfectedSymptoms<double>>(),
                        mb.parameters.get<mio::osecir::DynamicNPIsInfectedSymptoms<double>>()),
              "dynamic NPIs");
    }
    for (size_t e = 0; e < std::min(a.edges().size(), b.edges().size()); ++e) {
        const auto &x = a.edges()[e], &y = b.edges()[e];
        check(x.start_node_idx == y.start_node_idx && x.end_node_idx == y.end_node_idx, "edge ends");
        const auto &gx = x.property.get_coefficients(), &gy = y.property.get_coefficients();
        check(gx.get_num_matrices() == gy.get_num_matrices(), "coefficient vectors");
        for (size_t k = 0; k < std::min(gx.get_num_matrices(), gy.get_num_matrices()); ++k) {
            check(same_vector(gx[k].get_baseline(), gy[k].get_baseline()) &&
                      same_vector(gx[k].get_minimum(), gy[k].get_minimum()),
                  "edge baseline / minimum");
            check(gx[k].get_dampings().size() == gy[k].get_dampings().size(), "edge dampings");
            for (size_t d = 0; d < std::min(gx[k].get_dampings().size(), gy[k].get_dampings().size()); ++d) {
                const auto &p = gx[k].get_dampings()[d], &q = gy[k].get_dampings()[d];
                check(same_vector(p.coeffs, q.coeffs) && p.level == q.level && p.type == q.type && p.time == q.time,
                      "edge damping");
            }
        }
        check(same_npis(x.property.get_dynamic_npis_infected(), y.property.get_dynamic_npis_infected()), "edge NPIs");
    }
}

void write_text(const std::string& path, const std::string& text)
{
    std::ofstream(path) << text;
}

int errors(const std::string& dir)
{
    namespace fs = std::filesystem;
    using mio::StatusCode;
    auto r = mio::read_graph<double, mio::osecir::Model<double>>(dir + "/missing");
    check(!r && r.error().code() == StatusCode::FileNotFound, "missing directory");
    fs::create_directories(dir);
    auto g = example_graph();
    g.edges().erase(g.edges().begin(), g.edges().begin() + 3); // node 0 keeps no out-edge
    check(bool(mio::write_graph(g, dir)), "write graph without out-edges of node 0");
    r = mio::read_graph<double, mio::osecir::Model<double>>(dir);
    check(!r && r.error().code() == StatusCode::FileNotFound, "missing edge file");
    r = mio::read_graph<double, mio::osecir::Model<double>>(dir, mio::IOF_None, false);
    check(bool(r) && r.value().nodes().size() == 3 && r.value().edges().empty(), "nodes without edges");
    write_text(dir + "/GraphEdges_node0.json", "[{\"StartNodeIndex\": 0, \"EndNodeIndex\": 7, \"Parameters\": {}}]");
    r = mio::read_graph<double, mio::osecir::Model<double>>(dir);
    check(!r && r.error().code() == StatusCode::OutOfRange, "EndNodeIndex out of range");
    write_text(dir + "/GraphEdges_node0.json", "[{\"StartNodeIndex\": 0, \"EndNodeIndex\": 1, \"Parameters\": {}}]");
    r = mio::read_graph<double, mio::osecir::Model<double>>(dir);
    check(!r && r.error().code() == StatusCode::KeyNotFound, "missing Coefficients");
    write_text(dir + "/GraphNode1.json", "{\"NodeId\": \"one\", \"Model\": {}}");
    r = mio::read_graph<double, mio::osecir::Model<double>>(dir, mio::IOF_None, false);
    check(!r && r.error().code() == StatusCode::InvalidType, "NodeId not an integer");
    write_text(dir + "/GraphNode1.json", "{\"NodeId\": 1, \"Model\": ");
    r = mio::read_graph<double, mio::osecir::Model<double>>(dir, mio::IOF_None, false);
    check(!r && r.error().code() == StatusCode::InvalidFileFormat, "truncated file");
    return failures;
}

} // namespace

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "example" && argc == 3) {
        const Graph g  = example_graph();
        auto st        = mio::write_graph(g, argv[2]);
        check(bool(st), "write_graph: " + st.error().formatted_message());
        auto r = mio::read_graph<double, mio::osecir::Model<double>>(argv[2]);
        check(bool(r), "read_graph: " + r.error().formatted_message());
        if (r) {
            compare(g, r.value());
        }
    }
    else if (mode == "roundtrip" && argc == 4) {
        auto r = mio::read_graph<double, mio::osecir::Model<double>>(argv[2]);
        if (!r) {
            std::cout << "read_graph: " << r.error().formatted_message() << "\n";
            return 2;
        }
        auto st = mio::write_graph(r.value(), argv[3]);
        check(bool(st), "write_graph: " + st.error().formatted_message());
    }
    else if (mode == "errors" && argc == 3) {
        errors(argv[2]);
    }
    else {
        std::cout << "usage: test_mio_io example DIR | roundtrip IN OUT | errors DIR\n";
        return 2;
    }
    if (failures) {
        std::cout << failures << " FAILED\n";
        return 1;
    }
    std::cout << "ALL PASSED\n";
    return 0;
}
