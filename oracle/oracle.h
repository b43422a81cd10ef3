/*
 * oracle.h - CPU restatement of MEmilio's Graph-ODE path, used ONLY as the parity checker in tests/ and
 * as the CPU baseline leg of bench.py. Nothing in memilio_amd/ links or calls it.
 *
 * Every function follows the reference source line by line (citations in oracle.cpp). It keeps the
 * reference's loop structure and floating-point evaluation order (compiled with -ffp-contract=off), so
 * that it reproduces the reference's golden CSV fixtures to 1e-10.
 *
 * Inputs use the same plain descriptors as the device library (include/mgo.h).
 */
#ifndef MGO_ORACLE_H
#define MGO_ORACLE_H

#include "../include/mgo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* counters returned by the graph / ensemble runs */
typedef struct orc_stats {
    int64_t steps;           /* accepted (or forced) integration steps, summed over nodes */
    int64_t attempts;        /* try_step attempts (RK stage sets), summed over nodes */
    int64_t rhs_evals;       /* right-hand-side evaluations, nodes and edges */
    int64_t negative_fixups; /* compartments repaired by the return-underflow fix-up */
    int64_t returns;         /* traveler returns processed */
} orc_stats;

/* osecir::simulate(t0, tmax, cfg->dt_node, model, core): full TimeSeries of the single node of `model`.
 * Returns the number of time points (may exceed capacity; then only capacity points are written) or
 * a negative status. times[capacity], values[capacity][10*G]. */
int32_t orc_simulate(const mgo_graph_desc* model, const mgo_sim_cfg* cfg, double tmax, int32_t capacity,
                     double* times, double* values);

/* orc_simulate followed by interpolate_simulation_result (cpp/memilio/data/analyze_result.h:186-265):
 * daily_out[day][10*G]; returns the number of days (or negative status). */
int32_t orc_simulate_daily(const mgo_graph_desc* model, const mgo_sim_cfg* cfg, double tmax, int32_t capacity_days,
                           double* daily_out, orc_stats* stats);

/* make_mobility_sim(t0, dt_graph, graph).advance(tmax), then interpolate_simulation_result on every node.
 * daily_out[day][node][10*G], *num_days days from ceil(t0). Returns MGO_OK or an error. */
int32_t orc_graph_run(const mgo_graph_desc* graph, const mgo_sim_cfg* cfg, double tmax, int32_t capacity_days,
                      double* daily_out, int32_t* num_days, orc_stats* stats);

/* orc_graph_run with the node phase of every half-day on node_threads OpenMP threads (the CPU timing of a
 * single graph parallel over nodes; results equal orc_graph_run bit for bit). */
int32_t orc_graph_run_threads(const mgo_graph_desc* graph, const mgo_sim_cfg* cfg, double tmax, int32_t node_threads,
                              int32_t capacity_days, double* daily_out, int32_t* num_days, orc_stats* stats);

/* Same as orc_graph_run but also returns the full TimeSeries of one node (result of the reference's
 * get_result()); returns the number of points of that node (or negative status). */
int32_t orc_graph_run_node_series(const mgo_graph_desc* graph, const mgo_sim_cfg* cfg, double tmax, int32_t node,
                                  int32_t capacity, double* times, double* values);

/* ParameterStudy::run_serial over pre-sampled graphs, runs distributed over `threads` OpenMP threads
 * (each run serial, as run_impl). daily_out as mgo_ensemble_run (may be NULL). */
int32_t orc_ensemble_run(const mgo_ensemble_desc* desc, int32_t threads, double* daily_out, orc_stats* stats);

/* SEIR model nodes (cpp/models/ode_seir/model.h:47-102) used by the reference's graph fixture.
 * params[node][3*G] = TimeExposed[G], TimeInfected[G], TransmissionProbabilityOnContact[G];
 * contact[node][G][G]; populations[node][G][4] (S, E, I, R). Runs make_mobility_sim with the default
 * core and returns the full TimeSeries of node `node` (number of points, or negative status). */
int32_t orc_oseir_graph_series(int32_t num_nodes, int32_t G, const double* params, const double* contact,
                               const double* populations, int32_t num_edges, const int32_t* edge_start,
                               const int32_t* edge_end, const double* edge_coeffs, double t0, double dt_graph,
                               double dt_node, double tmax, int32_t node, int32_t capacity, double* times,
                               double* values);

/* Known-answer scenario of test_mobility.cpp:154-215: two SimulationNodes built from node 0 of `model` at
 * time t, one MobilityEdge with coefficients coeffs[10*G]; apply_mobility at t (departure), advance both
 * nodes by 0.5, apply at t+0.5 (return), advance, apply at t+1 (departure). out[6][10*G] = (node1, node2)
 * after each of the three edge calls. */
int32_t orc_edge_scenario(const mgo_graph_desc* model, double t, const double* coeffs, double* out);

/* One integrator step of a core on y' = f(t) test problems (test_numericalIntegration.cpp:82-247):
 * problem 0: y' = cos(t) (scalar). Runs SystemIntegrator::advance from t0 to tmax with dt, returns the
 * number of points (times/values as orc_simulate, values[capacity][1]). */
int32_t orc_integrate_test_problem(int32_t problem, const mgo_integrator_cfg* core, double t0, double tmax,
                                   double dt, int32_t capacity, double* times, double* values);

/* implement_dynamic_npis (cpp/memilio/epidemiology/dynamic_npis.h:318-392) on a group of damping lists
 * (known-answer scenarios of cpp/tests/test_dynamic_npis.cpp); see oracle.cpp for the argument layout. */
int32_t orc_implement_npis(int32_t rows, int32_t cols, int32_t num_mats, const int32_t* list_offsets,
                           const double* times, const int32_t* levels, const int32_t* types, const double* coeffs,
                           int32_t num_npis, const double* npi_values, const int32_t* npi_levels,
                           const int32_t* npi_types, const int32_t* npi_mat_offsets, const int32_t* npi_mats,
                           int32_t num_weights, const double* npi_weights, double begin, double end,
                           int32_t make_matrix, int32_t capacity, int32_t* out_offsets, double* out_times,
                           int32_t* out_levels, int32_t* out_types, double* out_coeffs);

/* osecir::Simulation::advance with dynamic NPIs over a constant-state simulation (the reference tests'
 * MockSimulation): contact(0,0) of node 0 at the query times afterwards, and its number of dampings. */
int32_t orc_npi_frozen_run(const mgo_graph_desc* model, double t0, double tmax, int32_t num_query,
                           const double* query_times, double* out_c00, int32_t* out_num_dampings);

/* Edge dynamic-NPI checks of MobilityEdge::apply_mobility over a sequence of calls (times, infections
 * relative of the source); number of dampings of the coefficient vector after each call. */
int32_t orc_edge_npi_calls(int32_t n, double threshold, double base, double duration, double value,
                           int32_t num_weights, const double* weights, int32_t num_calls, const double* times,
                           const double* inf_rel, int32_t* out_num_dampings);

/* Effective contact matrix entry of node 0 at time t (ContactMatrixGroup::get_matrix_at(t)(i, j)). */
double orc_contact_at(const mgo_graph_desc* model, double t, int32_t i, int32_t j);

#ifdef __cplusplus
}
#endif

#endif
