/*
 * mgo.h - C ABI of the MI355X Graph-ODE engine (age-stratified SECIR metapopulation model).
 *
 * Plain C: POD structs, raw pointers and sizes, integer status codes, no exceptions across the
 * boundary, caller-owned host buffers. One context per host thread, or serialise calls on a shared
 * context (every entry point except mgo_serve_* takes the context lock).
 *
 * Reference surface each entry point replaces (paths relative to the memilio source tree):
 *   mgo_graph_create + mgo_graph_advance
 *       == mio::make_mobility_sim(t0, dt, graph) + GraphSimulation::advance(tmax)
 *          cpp/memilio/mobility/metapopulation_mobility_instant.h:656-678,
 *          cpp/memilio/mobility/graph_simulation.h:101-120
 *   mgo_simulate_single
 *       == mio::osecir::simulate(t0, tmax, dt, model, core) + interpolate_simulation_result
 *          cpp/models/ode_secir/model.h:385-390, cpp/memilio/data/analyze_result.h:186-265
 *   mgo_ensemble_run
 *       == mio::ParameterStudy::run / run_serial over pre-sampled graphs
 *          cpp/memilio/compartments/parameter_studies.h:112-130,156-217,288-345,383-399
 *   mgo_serve_*  (no reference counterpart: continuous-batching request loop over the same engine)
 *
 * State vector layout of one node: index = g*10 + c, with c in the osecir::InfectionState order
 * S, E, INS, INSC, ISy, ISyC, ISev, ICr, R, D  (cpp/models/ode_secir/infection_state.h:33-47).
 */
#ifndef MGO_H
#define MGO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes
 * Mirrors mio::StatusCode (cpp/memilio/io/io.h:49) where a counterpart exists. */
typedef int32_t mgo_status;
#define MGO_OK 0
#define MGO_ERR_INVALID_VALUE 1   /* StatusCode::InvalidValue */
#define MGO_ERR_OUT_OF_RANGE 2    /* StatusCode::OutOfRange   */
#define MGO_ERR_OUT_OF_MEMORY 3   /* device or HBM-budget allocation failed */
#define MGO_ERR_UNSUPPORTED 4     /* valid for the reference, not supported by the device path */
#define MGO_ERR_DEVICE 5          /* HIP runtime error */
#define MGO_ERR_NOT_FOUND 6       /* unknown graph / batch / ticket id */
#define MGO_ERR_BUSY 7            /* serving queue full */

/* number of compartments per age group (osecir::InfectionState::Count) */
#define MGO_NUM_COMPARTMENTS 10
/* largest number of age groups the device kernels are built for */
#define MGO_MAX_GROUPS 8

/* ---------------------------------------------------------------- parameters
 * Per-node scalar parameters, index into mgo_graph_desc.node_scalars[node*MGO_NUM_SCALARS + k].
 * cpp/models/ode_secir/parameters.h:46-394 (tags) and :417-450 (commuter detection fields). */
enum {
    MGO_P_START_DAY = 0,                    /* StartDay, default 0 */
    MGO_P_SEASONALITY = 1,                  /* Seasonality, default 0 */
    MGO_P_ICU_CAPACITY = 2,                 /* ICUCapacity, default DBL_MAX */
    MGO_P_TEST_AND_TRACE_CAPACITY = 3,      /* TestAndTraceCapacity, default DBL_MAX */
    MGO_P_TEST_AND_TRACE_MAX_RISK = 4,      /* TestAndTraceCapacityMaxRisk, default 5 */
    MGO_P_COMMUTER_NONDETECTION = 5,        /* get_commuter_nondetection, default 0 */
    MGO_P_START_COMMUTER_DETECTION = 6,     /* default 0 */
    MGO_P_END_COMMUTER_DETECTION = 7,       /* default 0 */
    MGO_NUM_SCALARS = 8
};

/* Per-node, per-age-group parameters, index into
 * mgo_graph_desc.age_params[(node*MGO_NUM_AGE_PARAMS + k)*G + g]. */
enum {
    MGO_A_TIME_EXPOSED = 0,
    MGO_A_TIME_INFECTED_NO_SYMPTOMS = 1,
    MGO_A_TIME_INFECTED_SYMPTOMS = 2,
    MGO_A_TIME_INFECTED_SEVERE = 3,
    MGO_A_TIME_INFECTED_CRITICAL = 4,
    MGO_A_TRANSMISSION_PROBABILITY_ON_CONTACT = 5,
    MGO_A_RELATIVE_TRANSMISSION_NO_SYMPTOMS = 6,
    MGO_A_RECOVERED_PER_INFECTED_NO_SYMPTOMS = 7,
    MGO_A_RISK_OF_INFECTION_FROM_SYMPTOMATIC = 8,
    MGO_A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC = 9,
    MGO_A_SEVERE_PER_INFECTED_SYMPTOMS = 10,
    MGO_A_CRITICAL_PER_SEVERE = 11,
    MGO_A_DEATHS_PER_SEVERE = 12,
    MGO_A_DEATHS_PER_CRITICAL = 13,
    MGO_NUM_AGE_PARAMS = 14
};

/* ---------------------------------------------------------------- graph description
 * One Graph<SimulationNode<osecir::Simulation>, MobilityEdge> (cpp/memilio/mobility/graph.h:144-303)
 * in structure-of-arrays form. All pointers are host pointers and are only read during the call.
 *
 * Contact patterns: one ContactMatrix per node (ContactMatrixGroup of size 1,
 * cpp/memilio/epidemiology/contact_matrix.h:190-194,392-400): baseline B, minimum M (NULL = 0) and a
 * list of dampings (cpp/memilio/epidemiology/damping.h:73-140) given per node in CSR form.
 * The effective matrix is B - D(t) o (B - M) with D(t) the accumulated damping (damping.h:381-396).
 *
 * Edges: MobilityEdge coefficients, one vector of length 10*G per edge (metapopulation_mobility_instant.h:
 * 319-327), or one value per edge that applies to every compartment (edge_coeff_len = 1, the
 * Eigen::VectorX::Constant form the reference's examples use). Edges may be given in any order;
 * duplicates (same start,end) keep the last one, as Graph::add_edge's insert_sorted_replace does
 * (graph.h:231-241).
 */
/* Dynamic NPIs of the nodes' contact matrices (osecir DynamicNPIsInfectedSymptoms,
 * cpp/models/ode_secir/parameters.h:342-355; mio::DynamicNPIs, cpp/memilio/epidemiology/dynamic_npis.h:36-242).
 * Checked as osecir::Simulation::advance does (cpp/models/ode_secir/model.h:289-367): at the start of every
 * node advance and at the end of an active NPI, against InfectedSymptoms / total population * base value;
 * an exceeded threshold adds its dampings to the node's contact matrix (implement_dynamic_npis,
 * dynamic_npis.h:318-392, with make_contact_damping_matrix of value * group weights). */
enum {
    MGO_NPI_DURATION = 0,    /* DynamicNPIs::get_duration (default 14) */
    MGO_NPI_DELAY,           /* get_implementation_delay (default 0) */
    MGO_NPI_BASE_VALUE,      /* get_base_value (default 1) */
    MGO_NPI_DIRECTIVE_BEGIN, /* get_directive_begin (default lowest()) */
    MGO_NPI_DIRECTIVE_END,   /* get_directive_end (default max()) */
    MGO_NPI_NUM_PARAMS
};

typedef struct mgo_dynamic_npis {
    const int32_t* node_offsets;    /* [num_nodes+1]: thresholds of node v are [node_offsets[v], node_offsets[v+1]) */
    const double* node_params;      /* [num_nodes][MGO_NPI_NUM_PARAMS] */
    const double* thresholds;       /* [T] threshold values (set_threshold order; equal values: the later wins) */
    const int32_t* damping_offsets; /* [T+1]: dampings of threshold k are [damping_offsets[k], damping_offsets[k+1]) */
    const double* damping_values;   /* [D] DampingSampling value */
    const int32_t* damping_levels;  /* [D] DampingLevel */
    const int32_t* damping_types;   /* [D] DampingType */
    const double* damping_weights;  /* [D][G] group weights */
} mgo_dynamic_npis;

/* Device limit: entries of a node's damping list (its initial dampings plus those dynamic NPIs add) */
#define MGO_NPI_LIST_EXTRA 32

typedef struct mgo_graph_desc {
    int32_t num_nodes;
    int32_t num_groups;               /* G, 1..MGO_MAX_GROUPS */
    const double* node_scalars;       /* [num_nodes][MGO_NUM_SCALARS] */
    const double* age_params;         /* [num_nodes][MGO_NUM_AGE_PARAMS][G] */
    const double* contact_baseline;   /* [num_nodes][G][G] */
    const double* contact_minimum;    /* [num_nodes][G][G] or NULL (all zero) */
    const int32_t* damping_offsets;   /* [num_nodes+1] CSR offsets into the damping arrays, or NULL */
    const double* damping_times;      /* [K] SimulationTime of each damping */
    const int32_t* damping_levels;    /* [K] DampingLevel */
    const int32_t* damping_types;     /* [K] DampingType */
    const double* damping_coeffs;     /* [K][G][G] */
    const double* populations;        /* [num_nodes][G][10] initial compartment values */
    int32_t num_edges;
    const int32_t* edge_start;        /* [num_edges] */
    const int32_t* edge_end;          /* [num_edges] */
    const double* edge_coeffs;        /* [num_edges][edge_coeff_len] */
    int32_t edge_coeff_len;           /* 10*G, or 1 (one coefficient for all compartments); 0 means 10*G */
    const mgo_dynamic_npis* dynamic_npis; /* NULL = no dynamic NPIs */
    /* Mobility-coefficient dampings and edge dynamic NPIs (MobilityCoefficientGroup and
     * MobilityParameters::get_dynamic_npis_infected, cpp/memilio/mobility/metapopulation_mobility_instant.h:106-290;
     * applied by MobilityEdge::apply_mobility, :549-566 and :605-621: coefficients at time t are
     * B - D(t) (B - M)). Every out-edge of a node shares the node's out-edge damping list and NPIs, as
     * osecir::draw_sample gives every edge the same ones (cpp/models/ode_secir/parameter_space.h:240-246);
     * dampings are per age group (the value of group g applies to its 10 compartments, as
     * make_mobility_damping_vector builds them, damping_sampling.h:298-310). */
    const double* edge_coeff_minimum;     /* [num_edges][edge_coeff_len] minimum coefficients M, or NULL (zero) */
    const int32_t* out_damping_offsets;   /* [num_nodes+1] CSR into the out_damping arrays, or NULL */
    const double* out_damping_times;      /* [K] */
    const int32_t* out_damping_levels;    /* [K] */
    const int32_t* out_damping_types;     /* [K] */
    const double* out_damping_coeffs;     /* [K][G] damping value per age group */
    const mgo_dynamic_npis* out_dynamic_npis; /* per node: NPIs of its out-edges (node_params: duration and base
                                                 value are used; apply_mobility has no delay or directive), or NULL */
} mgo_graph_desc;

/* ---------------------------------------------------------------- integrators
 * cpp/memilio/math/stepper_wrapper.h:56-259, cpp/memilio/math/adapt_rk.h:145-310, euler.h:32-63 */
enum {
    /* ControlledStepperWrapper<runge_kutta_cash_karp54>: the default core of mio::Simulation
     * (cpp/memilio/compartments/simulation.h:33-34); Boost.odeint controller semantics. */
    MGO_INTEGRATOR_CASH_KARP54 = 0,
    /* RKIntegratorCore with the default Runge-Kutta-Fehlberg 4(5) tableau (adapt_rk.h:71-137). */
    MGO_INTEGRATOR_RKF45 = 1,
    /* ExplicitStepperWrapper<runge_kutta_cash_karp54>: fixed step, 5th-order solution propagated. */
    MGO_INTEGRATOR_CASH_KARP54_FIXED = 2,
    /* EulerIntegratorCore */
    MGO_INTEGRATOR_EULER = 3
};

typedef struct mgo_integrator_cfg {
    int32_t kind;      /* MGO_INTEGRATOR_* */
    double abs_tol;    /* adaptive cores only */
    double rel_tol;
    double dt_min;     /* adaptive cores only; the reference defaults are DBL_MIN and DBL_MAX */
    double dt_max;
} mgo_integrator_cfg;

/* Per-simulation settings. */
typedef struct mgo_sim_cfg {
    double t0;          /* start time of every node and of the graph simulation */
    double dt_graph;    /* GraphSimulation step (0.5 = daily exchange); ignored by mgo_simulate_single */
    double dt_node;     /* initial integration step of each node's Simulation (reference default 0.1) */
    double horizon;     /* largest tmax that will be requested; sizes the daily output buffer */
    mgo_integrator_cfg integrator;
} mgo_sim_cfg;

/* ---------------------------------------------------------------- context */
typedef struct mgo_context mgo_context;

typedef struct mgo_device_cfg {
    int32_t device;            /* HIP device ordinal */
    size_t hbm_budget_bytes;   /* 0 = 90% of free device memory at creation */
    int32_t flags;             /* MGO_CTX_* */
} mgo_device_cfg;

#define MGO_CTX_NO_GRAPH_CAPTURE 1   /* launch the half-day kernels directly instead of replaying a hipGraph */
#define MGO_CTX_PROFILE_KERNELS 2    /* time every node / edge kernel launch with HIP events (no graph capture) */
#define MGO_CTX_SEQUENTIAL_FIXUP 4   /* resolve return-underflow fix-ups with the sequential replay only */

mgo_status mgo_context_create(const mgo_device_cfg* cfg, mgo_context** out);
mgo_status mgo_context_destroy(mgo_context* ctx);
/* Replaces the MGO_CTX_* flags; they apply to every batch run and graph advance started afterwards. */
mgo_status mgo_context_set_flags(mgo_context* ctx, int32_t flags);
/* Layout of ensembles created afterwards: runs that share a graph (same edge arrays and node count) are
 * interleaved node-major in tiles of `tile` runs, so that a wavefront holds one graph node of several
 * runs. 0 (the default) = autotuned: the first ensemble of a shape times a short trial of candidate
 * tiles on a subset of its runs and keeps the fastest for that shape. Results do not depend on the tile. */
mgo_status mgo_context_set_tile(mgo_context* ctx, int32_t tile);
/* Last error message of the context (thread-local copy, valid until the next call on that context). */
const char* mgo_last_error(const mgo_context* ctx);
/* Version string of the library. */
const char* mgo_version(void);

/* ---------------------------------------------------------------- graph simulation (plain path)
 * make_mobility_sim(t0, dt, graph): uploads parameters and initial values; state stays on the device. */
typedef int64_t mgo_graph_id;
mgo_status mgo_graph_create(mgo_context* ctx, const mgo_graph_desc* graph, const mgo_sim_cfg* cfg,
                            mgo_graph_id* out);
/* GraphSimulation::advance(tmax). Returns after the device work is complete. */
mgo_status mgo_graph_advance(mgo_context* ctx, mgo_graph_id id, double tmax);
/* GraphSimulation::get_t() */
mgo_status mgo_graph_get_time(mgo_context* ctx, mgo_graph_id id, double* t);
/* Last value of every node's result TimeSeries, [num_nodes][10*G]. */
mgo_status mgo_graph_read_state(mgo_context* ctx, mgo_graph_id id, double* y_out);
/* Daily values as interpolate_simulation_result(graph) gives them: out[(day-day0)][node][10*G] for the
 * integer days day0..day0+*num_days-1 reached so far (day0 = ceil(t0)). capacity_days limits the copy. */
mgo_status mgo_graph_read_daily(mgo_context* ctx, mgo_graph_id id, int32_t capacity_days, double* out,
                                int32_t* num_days);
/* Per-node integrator statistics since creation: accepted steps, attempts (RHS-stage sets). [num_nodes] each.
 * Either pointer may be NULL. */
mgo_status mgo_graph_read_stats(mgo_context* ctx, mgo_graph_id id, int64_t* steps, int64_t* attempts);
mgo_status mgo_graph_destroy(mgo_context* ctx, mgo_graph_id id);

/* ---------------------------------------------------------------- single model
 * osecir::simulate(t0, tmax, dt, model, core) followed by interpolate_simulation_result: the graph must
 * have num_edges == 0; every node is an independent model. daily_out[(day-day0)][node][10*G],
 * *num_days set to the number of days written (capacity_days limits it). */
mgo_status mgo_simulate_single(mgo_context* ctx, const mgo_graph_desc* models, const mgo_sim_cfg* cfg,
                               double tmax, int32_t capacity_days, double* daily_out, int32_t* num_days);

/* ---------------------------------------------------------------- ensemble (ParameterStudy::run)
 * Runs num_runs pre-sampled graph simulations as one device batch. Graphs whose edge arrays are the same
 * pointers share one device copy of the edge structure. daily_out (may be NULL) receives, run after run,
 * [ndays_i][num_nodes_i][10*G_i] with ndays_i = floor(tmax_i) - ceil(t0_i) + 1. */
typedef struct mgo_ensemble_desc {
    int32_t num_runs;
    const mgo_graph_desc* graphs;   /* [num_runs] */
    const mgo_sim_cfg* cfgs;        /* [num_runs] */
    const double* tmax;             /* [num_runs] */
} mgo_ensemble_desc;

/* Timing and work counters of the last ensemble / serve step, filled if non-NULL. */
typedef struct mgo_run_info {
    double device_ms;          /* time between the first and last device event of the run */
    double node_days;          /* sum over runs of nodes * (tmax - t0) */
    int64_t steps;             /* accepted integration steps summed over all nodes */
    int64_t attempts;          /* step attempts summed over all nodes */
    int64_t exact_replays;     /* return phases resolved by the sequential negative-value fix-up replay */
    int64_t fixup_passes;      /* extra parallel return passes run to settle negative-value fix-ups */
    int32_t half_day_cycles;   /* graph steps executed */
    int32_t node_launches;     /* node-phase kernel launches timed (MGO_CTX_PROFILE_KERNELS) */
    double node_kernel_ms;     /* summed duration of those launches */
    int32_t edge_launches;     /* edge-phase kernel launches timed */
    double edge_kernel_ms;
    int32_t tile;              /* runs per node-major tile used for runs that share a graph */
    int32_t reserved;
} mgo_run_info;

mgo_status mgo_ensemble_run(mgo_context* ctx, const mgo_ensemble_desc* desc, double* daily_out,
                            mgo_run_info* info);

/* Batched form with device-resident results, for callers that post-process on the device or time the
 * engine without the download: create, advance all runs to their tmax, read back, destroy. */
typedef int64_t mgo_batch_id;
mgo_status mgo_batch_create(mgo_context* ctx, const mgo_ensemble_desc* desc, mgo_batch_id* out);
mgo_status mgo_batch_run(mgo_context* ctx, mgo_batch_id id, mgo_run_info* info);
/* Daily values of runs first_run .. first_run+num_runs-1, laid out as mgo_ensemble_run's daily_out. */
mgo_status mgo_batch_read_daily(mgo_context* ctx, mgo_batch_id id, int32_t first_run, int32_t num_runs,
                                double* daily_out);
/* Restores the batch to its state at creation (initial values, clocks, step sizes, counters), so that the
 * same resident inputs can be run again. */
mgo_status mgo_batch_reset(mgo_context* ctx, mgo_batch_id id);
mgo_status mgo_batch_destroy(mgo_context* ctx, mgo_batch_id id);

/* Ensemble statistics over runs first_run .. first_run+num_runs-1 of a batch, computed on the device from
 * the resident daily outputs. The runs must have the same shape (age groups, integrator, node count,
 * output days). Layout of a statistic: [day][node][10*G] like one run's daily output.
 *   mgo_batch_ensemble_mean        == ensemble_mean (cpp/memilio/data/analyze_result.h:289-315): sum over
 *                                     runs, in run order, of value / num_runs
 *   mgo_batch_ensemble_percentiles == ensemble_percentile (analyze_result.h:316-343) for each p[q] in
 *                                     (0, 1): the value at index floor(num_runs * p) of the sorted run values;
 *                                     out[q][day][node][10*G]; at most 8192 runs
 *   mgo_batch_sum_nodes            == sum_nodes (analyze_result.h:267-287): out[run][day][10*G], the nodes
 *                                     summed in node order */
mgo_status mgo_batch_ensemble_mean(mgo_context* ctx, mgo_batch_id id, int32_t first_run, int32_t num_runs,
                                   double* mean_out);
mgo_status mgo_batch_ensemble_percentiles(mgo_context* ctx, mgo_batch_id id, int32_t first_run, int32_t num_runs,
                                          const double* p, int32_t num_p, double* out);
mgo_status mgo_batch_sum_nodes(mgo_context* ctx, mgo_batch_id id, int32_t first_run, int32_t num_runs, double* out);

/* The same statistics over results held in host memory (ensemble_mean / ensemble_percentile of
 * cpp/memilio/data/analyze_result.h:289-343 applied to time series the caller holds, e.g. the interpolated
 * results of a parameter study): values[run][position] with num_positions positions per run (for the
 * reference's [run][node] time series: position = (node * num_time_points + time) * num_elements + element).
 * The values are copied to the device and reduced there by the kernels of mgo_batch_ensemble_*; results
 * have one run's layout: mean_out[position], out[q][position]. Percentiles: each p[q] in (0, 1), at most
 * 8192 runs. */
mgo_status mgo_ensemble_mean_values(mgo_context* ctx, const double* values, int32_t num_runs, int64_t num_positions,
                                    double* mean_out);
mgo_status mgo_ensemble_percentiles_values(mgo_context* ctx, const double* values, int32_t num_runs,
                                           int64_t num_positions, const double* p, int32_t num_p, double* out);

/* Allocation-free form of mgo_ensemble_percentiles_values for values already in device memory (no counterpart
 * in the reference; for callers that cannot allocate: code under HIP graph capture, tight serving loops).
 * values_dev[run][position] and out_dev[q][position] are device pointers of the caller; the per-run offsets and
 * ranks the kernels read live in a fixed device scratch of the library (one per GPU, held until the process
 * ends): at most 8192 runs and 64 percentiles, else MGO_ERR_OUT_OF_RANGE; no device or host memory is
 * allocated. stream (a hipStream_t) = NULL: the call runs on the context's stream and returns when out_dev is
 * complete; else it is enqueued on that stream and returns at once. Calls overlapping on one GPU (several
 * streams, contexts or threads) share the scratch and race: order them on one stream or synchronise. */
mgo_status mgo_ensemble_percentiles_noalloc(mgo_context* ctx, const double* values_dev, int32_t num_runs,
                                            int64_t num_positions, const double* p, int32_t num_p, double* out_dev,
                                            void* stream);
/* sum_nodes (analyze_result.h:267-287) of host values[run][day][node][element] (num_elements per node):
 * out[run][day][element], the nodes summed in node order, on the device (the kernel of mgo_batch_sum_nodes). */
mgo_status mgo_sum_nodes_values(mgo_context* ctx, const double* values, int32_t num_runs, int32_t num_days,
                                int32_t num_nodes, int32_t num_elements, double* out);

/* ---------------------------------------------------------------- serving loop
 * A worker thread owned by the context runs half-day cycles over the resident requests; requests are
 * admitted as HBM budget and slots free up (continuous batching), so requests submitted while others
 * run overlap with them on the device. Each request's arithmetic is independent of the others, so a
 * result is the same alone and in a mix. */
typedef int64_t mgo_ticket;

typedef struct mgo_request {
    mgo_graph_desc graph;
    mgo_sim_cfg cfg;
    double tmax;
} mgo_request;

typedef struct mgo_serve_cfg {
    int32_t max_resident;        /* upper bound on concurrently resident requests (0 = 4096) */
    size_t hbm_budget_bytes;     /* 0 = the context budget, at most 32 GiB */
    int32_t flags;               /* MGO_SERVE_* */
    size_t result_pool_bytes;    /* pinned host memory holding finished requests' daily values until
                                    mgo_serve_take (0 = 2 GiB); a result that does not fit is staged in
                                    pageable memory instead (slower copy) */
} mgo_serve_cfg;

/* Serving launches one set of kernels per integrator kind over the resident requests of every age-group count
 * (blocks dispatch on their segment's G). With MGO_SERVE_BATCH_KERNELS it launches one set per (integrator,
 * G) instead: the very kernels a batch or graph run uses, so a served request's results are bitwise equal to
 * those runs; more launches per simulated day (higher latency at low load). Either way a request's result
 * does not depend on the other requests. */
#define MGO_SERVE_BATCH_KERNELS 1

/* Scenario update on a resident request: contact-matrix scaling s applied from time t_start as a damping
 * of coefficient 1-s on every node (ContactMatrix::add_damping semantics, damping.h:381-396,616-658). If
 * t_start < 0 the update starts one day after the request's current simulation time. */
typedef struct mgo_damping_update {
    double scale;        /* s in (0, 1] */
    double t_start;      /* damping time */
    int32_t level;
    int32_t type;
} mgo_damping_update;

typedef struct mgo_result_info {
    mgo_ticket ticket;
    mgo_status status;          /* MGO_OK or the error that ended the request. MGO_ERR_OUT_OF_MEMORY with
                                   num_days > 0: the request ran to its end, but a scenario update that did
                                   not fit the serving arena was dropped (mgo_serve_updates_applied lists the
                                   updates it ran with); the daily values are those of that run */
    double submit_ms;           /* context clock at submission */
    double admit_ms;            /* context clock at admission to the device */
    double retire_ms;           /* context clock when its last day was computed and its slot freed */
    double done_ms;             /* context clock when its daily values were in host memory */
    int32_t num_nodes;
    int32_t num_groups;
    int32_t num_days;           /* daily values available */
} mgo_result_info;

mgo_status mgo_serve_start(mgo_context* ctx, const mgo_serve_cfg* cfg);
mgo_status mgo_serve_submit(mgo_context* ctx, const mgo_request* req, mgo_ticket* out);
mgo_status mgo_serve_update(mgo_context* ctx, mgo_ticket ticket, const mgo_damping_update* upd);
/* Non-blocking: fills up to capacity finished requests; *count set to the number filled. */
mgo_status mgo_serve_poll(mgo_context* ctx, mgo_result_info* out, int32_t capacity, int32_t* count);
/* Daily values of a finished request, [day][node][10*G]; releases the request's host copy. */
mgo_status mgo_serve_take(mgo_context* ctx, mgo_ticket ticket, double* daily_out, size_t capacity_doubles);
/* Scenario updates a finished request ran with, in the order they were applied, each with the damping
 * time it was given (t_start resolved for updates submitted with t_start < 0); *count = their number,
 * up to capacity of them written. Call before mgo_serve_take. MGO_ERR_BUSY while the request runs. Running
 * the same request alone with these updates (t_start as reported) gives its result bit for bit.
 * (No counterpart in the reference: reports what the serving loop decided.) */
mgo_status mgo_serve_updates_applied(mgo_context* ctx, mgo_ticket ticket, mgo_damping_update* out,
                                     int32_t capacity, int32_t* count);
/* Waits until every submitted request is finished, then stops the worker. */
mgo_status mgo_serve_stop(mgo_context* ctx);

/* Counters of the serving loop since mgo_serve_start (monitoring; no counterpart in the reference). */
typedef struct mgo_serve_stats {
    int64_t iterations;        /* worker loop iterations (admission / retirement points) */
    int64_t cycles;            /* half-day cycles launched */
    int64_t admitted;
    int64_t retired;
    double node_cycles;        /* sum over cycles of the resident nodes advanced */
    double device_ms;          /* worker time from an iteration's first launch to the end of its cycles */
    double host_ms;            /* worker time in admission, updates, launch-table rebuilds and retirement */
    int32_t max_resident_seen; /* most requests resident at once */
    int32_t reserved;
} mgo_serve_stats;
mgo_status mgo_serve_get_stats(mgo_context* ctx, mgo_serve_stats* out);
/* Milliseconds on the context clock (steady clock, zero at context creation). */
double mgo_clock_ms(const mgo_context* ctx);

#ifdef __cplusplus
}
#endif

#endif /* MGO_H */
