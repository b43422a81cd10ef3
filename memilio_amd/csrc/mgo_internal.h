// mgo_internal.h - data layout shared by the host engine (mgo_host.cpp) and the HIP kernels
// (mgo_kernels.hip). Not part of the public ABI.
//
// Layout in HBM. A *segment* is a contiguous node range with its own buffers: one whole batch class
// (all runs with the same number of age groups G and integrator kind) or one served request. A
// launch covers the segments of one class through a block -> segment table; kernels are specialised
// per (G, integrator). Within a segment:
//   lane slot   = node * G + g   (one GPU lane per (node, age group); 64/G nodes per wavefront)
//   y, ypre             [node][g][10]          state, pre-return state
//   yint                [node][g][YINT_STRIDE] state before the last step of a phase ending on an output day
//   snap                [node][g][SNAP_STRIDE] departure row: snapshot (10), ISy mobility factor, commuter
//                                              nondetection factor at the departure time
//   lcoef               [node][g][NUM_LC]      per-lane rate coefficients (see LaneCoef)
//   cbase, cdelta       [node][g][G]           contact row B_g., B_g. - M_g.
//   rc                  [node][g][RC_STRIDE]   return coefficients of the node as a destination, frozen at
//                                              the departure time (see RetCoef)
//   nodep               [node]                 scalar parameters + instance index
//   tnode, dtnode       [node]                 integrator time and carried step size
// Instances (one graph simulation each) live in a device table; a clock kernel turns the instance
// clock into a per-cycle record read by the node and edge kernels.
//
// Tiling. Runs of one class that share a graph (same edge arrays and node count, e.g. the samples of a
// parameter study) can be interleaved node-major in tiles of T runs: local node v of the j-th run of a
// tile sits at segment node tile_base + v * T + j (node_base = tile_base + j, node_stride = T). A
// wavefront then holds the same graph node of 64/G consecutive runs: similar parameters (similar step
// counts under adaptive control) and identical edge lists (uniform edge walks, adjacent gathers).
// T = 1 is the run-major layout. Results do not depend on T. The node kernel (compute-bound, where step-
// count divergence matters) visits nodes through SegDev::node_order, which puts one graph node of all runs
// sharing the graph into consecutive lanes whatever T is; the edge kernels (gather-bound, where the
// working set of the gathered rows matters) follow the memory layout, so T is tuned for them.
#pragma once

#include <cstdint>

namespace mgo
{

constexpr int NC    = 10; // compartments per age group
constexpr int WAVE  = 64;
constexpr int BLOCK = 256;

enum Compartment : int { S = 0, E, INS, INSC, ISY, ISYC, ISEV, ICR, R, D };

// Per-lane coefficients. Each is the exact subexpression the reference multiplies the compartment
// with (cpp/models/ode_secir/model.h:154-212), so coef * y rounds as the reference's expression does.
enum LaneCoef : int {
    LC_E = 0,      // (1.0 / T_E)
    LC_INS_ISY,    // (1 - R) * (1 / T_INS)          (also INSC -> ISyC)
    LC_INS_R,      // R * (1.0 / T_INS)              (also INSC -> R)
    LC_ISY_ISEV,   // sev / T_ISy                    (also ISyC -> ISev)
    LC_ISY_R,      // (1 - sev) / T_ISy              (also ISyC -> R)
    LC_ISEV_ICR,   // cps / T_ISev                   (ICU capacity inactive)
    LC_ISEV_R,     // (1.0 - cps - dps) / T_ISev
    LC_ISEV_D,     // (dps + (cps - cps)) / T_ISev   (ICU capacity inactive)
    LC_ICR_D,      // dpc / T_ICr
    LC_ICR_R,      // (1.0 - dpc) / T_ICr
    LC_TP,         // TransmissionProbabilityOnContact
    LC_REL,        // RelativeTransmissionNoSymptoms
    LC_RISK,       // RiskOfInfectionFromSymptomatic
    LC_MAXRISK,    // MaxRiskOfInfectionFromSymptomatic
    LC_TTR,        // (1 - R) / T_INS   (test-and-trace demand per INS)
    LC_CPS,        // CriticalPerSevere
    LC_DPS,        // DeathsPerSevere
    LC_TISEV,      // TimeInfectedSevere
    NUM_LC
};

// Return coefficients of a destination lane, computed from its post-departure state (the `pop` of
// calculate_mobility_returns, metapopulation_mobility_instant.h:460-471) at the departure time.
enum RetCoef : int {
    RC_KAPPA = 0, // season(t_dep) * tp * sum_j C_gj(t_dep) w_j(pop): S->E rate per traveler in S
    RC_E,
    RC_INS_ISY,
    RC_INS_R,
    RC_ISY_ISEV,
    RC_ISY_R,
    RC_ISEV_ICR,  // ICU-adjusted at the departure state
    RC_ISEV_R,
    RC_ISEV_D,    // ICU-adjusted at the departure state
    RC_ICR_D,
    RC_ICR_R,
    RC_STRIDE = 12
};

// departure row of a lane: the snapshot of SimulationNode::m_last_state, the ISy mobility factor
// (get_mobility_factors) and the commuter nondetection factor (test_commuters) at the departure time;
// 96 bytes, read with 16-byte loads
constexpr int SNAP_FAC    = 10;
constexpr int SNAP_ND     = 11;
constexpr int SNAP_STRIDE = 12;

// last step of a lane's phase that ends on an output day d: the state before that step and the weight
// (d - t_prev) / (t_end - t_prev) of interpolate_simulation_result (analyze_result.h:235-252), or -1 when the
// step ends before d (the next phase's first step encloses d)
constexpr int YINT_W      = 10;
constexpr int YINT_STRIDE = 12;

// node flags
enum : int32_t {
    NF_TTC_ACTIVE  = 1, // TestAndTraceCapacity < DBL_MAX: the risk smoother depends on the state
    NF_ICU_ACTIVE  = 2, // ICUCapacity < DBL_MAX: the critical-care smoother depends on the state
    NF_DAMPED      = 4, // contact matrix has dampings
    NF_DETECTION   = 8, // commuter detection interval non-empty
    NF_NPI         = 16, // dynamic NPIs: the damping list lives in an NpiNode and changes during the run
    NF_OUT_DAMP    = 32, // out-edge coefficient dampings or edge NPIs (SegDev::out_npi)
};

// Dynamic NPIs of one node (osecir::Simulation::advance, models/ode_secir/model.h:289-367). The damping
// list, its cache (NodeP::damp_time / damp_mat point into it) and the NPI state change during the run;
// the kernel that changes them is the node kernel (one lane per node, see mgo_npi.h).
struct NpiNode {
    double duration, delay, base, begin, end; // DynamicNPIs parameters
    double total;    // the model's total population (get_infections_relative's denominator)
    double value;    // threshold of the active NPI (m_dynamic_npi.first), lowest() before the first
    double t_end;    // end of the active NPI (m_dynamic_npi.second)
    double* list;    // [list_cap][3 + G*G]
    double* cache_t; // [list_cap + 2]
    double* cache_m; // [list_cap + 2][G][G]
    double* work;    // npi_work_size(G, list_cap) doubles
    double* y0;      // [G][10] state at the start of the current node phase (return coefficients)
    int32_t thr_first, thr_count; // thresholds in SegDev::npi_thr (descending values)
    int32_t list_count, list_cap;
    int32_t cache_count, pad;
};

struct NodeP {
    double start_day;
    double seasonality;
    double ttc;       // TestAndTraceCapacity
    double ttc_hi;    // TestAndTraceCapacity * TestAndTraceCapacityMaxRisk
    double icu;       // ICUCapacity
    double nondetect; // commuter nondetection
    double det_start;
    double det_end;
    const double* damp_time; // damping cache: entry times (sentinels lowest()/max() included)
    const double* damp_mat;  // damping cache: [entry][G][G]
    int32_t flags;
    int32_t inst;            // instance index
    int32_t damp_cnt;        // number of damping-cache entries
    int32_t local;           // node index within its instance (graph node id order)
};

// per-instance state; the c_* fields are the record of the current half-day cycle
struct InstDev {
    double t;         // GraphSimulation::get_t()
    double dt_graph;
    double tmax;
    double t_dep;     // pending departure time
    double ret_time;  // pending return time
    double abs_tol, rel_tol, dt_min, dt_max;
    double c_t0, c_t1, c_step, c_tdep;
    double* out;      // daily values [day][node][G][10]
    int32_t node_base;
    int32_t num_nodes;
    int32_t estruct;  // edge structure index, -1 = no edges
    int32_t pending;
    int32_t toggle;   // MobilityEdge::m_return_mobile_population (identical on all edges)
    int32_t status;
    int32_t day0;
    int32_t ndays;
    int32_t c_active;
    int32_t c_action; // 0 none, 1 departure, 2 return
    int32_t c_out_day;
    int32_t replays;   // return phases resolved by the sequential replay so far
    int32_t exact_end; // single-model run: the last step ends exactly at tmax
    // negative-value fix-up iteration of the current return phase (see return_walk)
    int64_t fx_base;      // first in-edge record pair of this instance in SegDev::fx_rec (in pairs)
    int64_t fx_pool_base; // first double of this instance's slots in SegDev::fx_pool
    int32_t fx_pass;      // pass being run (0 = the first return walk)
    int32_t fx_state;     // 0 first pass, 1 iterating, 2 sequential replay needed, 3 settled
    int32_t fx_changed;   // set by a pass in which a fix-up record changed
    int32_t fx_count;     // pool slots used in this return phase
    int32_t fx_passes;    // extra passes run so far (statistics)
    int32_t fx_sequential; // resolve fix-ups with the sequential replay only
    int32_t seg;          // segment (index into LaunchDev::segs of its class) holding the nodes
    int32_t fx_epoch;     // record tag of pass 0 of the current return phase (grows by FX_EPOCH_STEP)
    int32_t node_stride;  // segment index of local node v = node_base + v * node_stride (see Tiling)
    int32_t pad3;
};

// Fix-up records. Each in-edge keeps two tagged versions {tag, slot}; tag = fx_epoch + pass of the pass
// that wrote it, slot = pool slot of the fixed traveler vector or -1 (unfixed). A pass p reads the newest
// version tagged below fx_epoch + p and overwrites the older one, so readers of the previous pass's
// records never race with writers of the current pass. Tags of earlier return phases are below fx_epoch
// and read as "no record" (unfixed).
struct alignas(8) FxRec {
    int32_t tag;
    int32_t slot;
};
constexpr int FX_SLOTS      = 1024; // pool slots per instance and return phase
constexpr int FX_MAX_PASSES = 8;    // passes before falling back to the sequential replay
constexpr int FX_EPOCH_STEP = 16;   // > FX_MAX_PASSES

// Cycle flags, written by the clock kernel (cleared, then set by the instances) and by fixup_check: which
// phases of this half-day cycle have work anywhere in the batch, so that the launches of idle phases (the
// departure kernel in a return cycle, fix-up passes nobody needs) exit at once.
constexpr int CYC_DEPART = 0;                    // some instance departs
constexpr int CYC_RETURN = 1;                    // some instance returns
constexpr int CYC_PASS0  = 2;                    // CYC_PASS0 + p: some instance runs fix-up pass p
constexpr int CYC_FLAGS  = CYC_PASS0 + FX_MAX_PASSES;

// CSR edge structure of one graph (node indices local to the instance). In-edges of node v are sorted
// by source; in_split[v] = number of in-edges whose source < v, so walking
//   in[0, split) , out[all] , in[split, end)
// visits v's edges in the reference's global (start, end) order (graph.h:231-241).
// one CSR entry: neighbour, coefficient (uniform) or index of its coefficient vector
struct EdgeEnt {
    double cval;      // coefficient if the edge's vector is uniform
    int32_t nbr;      // neighbour (local index)
    int32_t cidx;     // -1 uniform, else index of the [10*G] vector in cvec
    int32_t peer;     // out-entry: index of the same edge among the in-entries (its fix-up record)
    int32_t pad;
};

struct EdgeStructDev {
    const int32_t* in_ptr;   // [n+1]
    const int32_t* out_ptr;  // [n+1]
    const int32_t* split;    // [n]
    const int32_t* seq;      // [num_edges][3] = (start, end, out-entry) in global edge order, exact replay
    const EdgeEnt* in_ent;
    const EdgeEnt* out_ent;
    const double* cvec;      // non-uniform coefficient vectors [k][10*G]
    const double* in_min;    // minimum coefficient of each uniform in-entry, or NULL (zero)
    const double* out_min;   // minimum coefficient of each uniform out-entry, or NULL (zero)
    const double* cvec_min;  // minimum coefficient vectors [k][10*G] (parallel to cvec), or NULL (zero)
    int32_t n;
    int32_t num_edges;
};

struct SegDev {
    int32_t G;
    int32_t num_nodes;
    double* y;
    double* snap;
    double* ypre;
    double* rc;
    double* yint;                  // [node][g][YINT_STRIDE] last step before a phase-end output day
    const double* lcoef;
    const double* cbase;
    const double* cdelta;
    const NodeP* nodep;
    double* tnode;
    double* dtnode;
    uint32_t* steps;
    uint32_t* attempts;
    const EdgeStructDev* estructs; // indexed by InstDev::estruct
    const int32_t* node_order;     // node-kernel lane slot -> segment node, or NULL (identity); groups the
                                   // same graph node of the runs sharing a graph whatever the tile size
    FxRec* fx_rec;                 // [in-edge][2] tagged fix-up record versions
    double* fx_pool;               // fixed traveler vectors [slot][G][10]
    int32_t* fx_dirty;             // [2][node] tag of the pass that changed a record of the node's out-edges
                                   // (plane = pass parity)
    NpiNode* npi;                  // [node] dynamic NPIs (nodes with NF_NPI), or NULL
    const double* npi_thr;         // threshold values
    const int32_t* npi_thr_damp;   // [threshold + 1] first damping of each threshold in npi_damp
    const double* npi_damp;        // [damping][3 + G] value, level, type, group weights
    // out-edge coefficient dampings and edge NPIs (nodes with NF_OUT_DAMP): per-group damping lists (entries of
    // G values) shared by a node's out-edges, their thresholds and dampings as above, and the damping factor
    // D of every lane at the last departure (coefficients B - D (B - M)); out_d is NULL without out-dampings
    NpiNode* out_npi;
    const double* out_thr;
    const int32_t* out_thr_damp;
    const double* out_damp;
    double* out_d;                 // [node][g]
};

// One kernel launch over the segments of a class.
struct LaunchDev {
    const SegDev* segs;
    const int32_t* block_seg;   // [num_blocks] segment of each (XCD-remapped) block, -1 = padding
    const int32_t* block_first; // [num segments] first block of each segment
    const int32_t* inst_ids;    // instances of this class
    int32_t num_inst;
    int32_t num_blocks;         // launched blocks (multiple of 8)
    int32_t G;                  // age groups of every segment, or 0: segments differ (mixed kernels)
    int32_t kind;
    int32_t npi;                // some node has dynamic NPIs (per-G kernels only)
    int32_t pad;
    int32_t* cyc;               // [CYC_FLAGS] cycle flags of the batch (or NULL: every phase runs its checks)
};

} // namespace mgo
