// mgo_host.cpp - host engine behind the C ABI of include/mgo.h.
//
// Builds device-resident batches of graph simulations (parameters, state, CSR edge structures,
// damping caches, daily output buffers), drives the half-day cycle of
// GraphSimulation::advance (cpp/memilio/mobility/graph_simulation.h:101-120) with the HIP kernels of
// mgo_kernels.hip, and runs the serving loop (continuous batching over the same engine).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/mgo.h"
#include "mgo_internal.h"
#include "mgo_npi.h"

namespace mgo
{
hipError_t launch_clock(InstDev* inst, int n, int32_t* cyc, hipStream_t s);
hipError_t launch_scatter_slots(InstDev* table, const InstDev* src, const int32_t* ids, int n, hipStream_t s);
hipError_t launch_set_segments(InstDev* table, const int32_t* ids, int n, hipStream_t s);
hipError_t launch_node(const LaunchDev& c, InstDev* inst, hipStream_t s);
hipError_t launch_edge(const LaunchDev& c, InstDev* inst, hipStream_t s);
hipError_t launch_ensemble_mean(const double* out, const int64_t* base, int runs, int64_t positions, double* mean,
                                hipStream_t s);
hipError_t launch_sum_nodes(const double* out, const int64_t* base, int runs, int days, int nodes, int width,
                            double* sums, hipStream_t s);
hipError_t launch_ensemble_select(const double* out, const int64_t* base, int runs, int64_t positions,
                                  const int32_t* ranks, int num_ranks, double* sel, hipStream_t s);
} // namespace mgo

using namespace mgo;

namespace
{

struct Error {
    mgo_status code;
    std::string msg;
};

[[noreturn]] void fail(mgo_status code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error{code, buf};
}

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) {
            fail(MGO_ERR_OUT_OF_MEMORY, "%s: %s", what, hipGetErrorString(e));
        }
        fail(MGO_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    }
}

// First-fit sub-allocator over one device allocation (serving: per-request buffers without hipMalloc) or
// one pinned host allocation (serving: result staging); alloc / free are thread-safe.
struct Arena {
    char* base  = nullptr;
    size_t size = 0;
    size_t used = 0;
    bool host   = false;
    std::map<size_t, size_t> free_; // offset -> length
    std::mutex mu;
    void init(size_t bytes, bool pinned_host = false)
    {
        size = bytes;
        host = pinned_host;
        if (host) {
            hip_check(hipHostMalloc(reinterpret_cast<void**>(&base), size, hipHostMallocDefault), "pinned pool");
        }
        else {
            hip_check(hipMalloc(reinterpret_cast<void**>(&base), size), "serving arena");
        }
        free_.clear();
        free_[0] = size;
        used     = 0;
    }
    void release()
    {
        if (base) {
            if (host) {
                (void)hipHostFree(base);
            }
            else {
                (void)hipFree(base);
            }
        }
        base = nullptr;
    }
    // nullptr when no free range is large enough
    void* try_alloc(size_t bytes)
    {
        std::lock_guard<std::mutex> lk(mu);
        bytes = (bytes + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > bytes) {
                    free_[off + bytes] = len - bytes;
                }
                used += bytes;
                return base + off;
            }
        }
        return nullptr;
    }
    void* alloc(size_t bytes)
    {
        if (void* p = try_alloc(bytes)) {
            return p;
        }
        fail(MGO_ERR_OUT_OF_MEMORY, "serving arena exhausted (%zu of %zu bytes used)", used, size);
    }
    void free(void* p, size_t bytes)
    {
        std::lock_guard<std::mutex> lk(mu);
        bytes      = (bytes + 255) & ~size_t(255);
        size_t off = size_t(static_cast<char*>(p) - base);
        used -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
};

// allocations of the calling thread go to this arena when set (the serving worker)
thread_local Arena* tl_arena = nullptr;

// device buffer owned by a batch
struct DevBuf {
    void* p      = nullptr;
    size_t bytes = 0;
    Arena* arena = nullptr;
    DevBuf()     = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf()
    {
        release();
    }
    void release()
    {
        if (p) {
            if (arena) {
                arena->free(p, bytes);
            }
            else {
                (void)hipFree(p);
            }
        }
        p     = nullptr;
        bytes = 0;
        arena = nullptr;
    }
    void alloc(size_t b)
    {
        release(); // a second alloc / upload on the same buffer replaces it instead of leaking it
        bytes = std::max<size_t>(b, 8);
        arena = tl_arena;
        if (arena) {
            p = arena->alloc(bytes);
        }
        else {
            hip_check(hipMalloc(&p, bytes), "hipMalloc");
        }
    }
    template <class T>
    void upload(const std::vector<T>& v, hipStream_t s)
    {
        alloc(v.size() * sizeof(T));
        if (!v.empty()) {
            hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
        }
    }
    template <class T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
};

// ------------------------------------------------------------------------------------------ dampings
// Accumulated damping cache of one contact matrix (cpp/memilio/epidemiology/damping.h:549-658):
// entries (time, D) with sentinels at lowest() and max(); same-level dampings add, different levels
// combine as 1 - prod(1 - d), equal (time, type, level) replaces.
struct DampIn {
    double time;
    int level, type;
    std::vector<double> m;
};

void damping_cache(std::vector<DampIn> list, int G, std::vector<double>& times, std::vector<double>& mats)
{
    // insert_sorted_replace on key (time, type, level) in input order (utils/stl_util.h:82-96)
    std::vector<DampIn> sorted;
    auto less = [](const DampIn& a, const DampIn& b) {
        return std::make_tuple(a.time, a.type, a.level) < std::make_tuple(b.time, b.type, b.level);
    };
    for (auto& d : list) {
        auto lb = std::lower_bound(sorted.begin(), sorted.end(), d, less);
        if (lb != sorted.end() && !less(d, *lb)) {
            *lb = d;
        }
        else {
            sorted.insert(lb, d);
        }
    }
    const size_t n = size_t(G) * G;
    times.assign(1, std::numeric_limits<double>::lowest());
    mats.assign(n, 0.0);
    std::vector<std::tuple<std::vector<double>, int, int>> active;
    std::vector<std::pair<std::vector<double>, int>> by_level;
    for (auto& d : sorted) {
        auto ita = std::find_if(active.begin(), active.end(), [&](auto& a) {
            return std::get<1>(a) == d.level && std::get<2>(a) == d.type;
        });
        if (ita != active.end()) {
            auto& sum = *std::find_if(by_level.begin(), by_level.end(), [&](auto& s) {
                return s.second == d.level;
            });
            auto& act = std::get<0>(*ita);
            for (size_t k = 0; k < n; ++k) {
                sum.first[k] += d.m[k] - act[k];
            }
            act = d.m;
        }
        else {
            active.emplace_back(d.m, d.level, d.type);
            auto its = std::find_if(by_level.begin(), by_level.end(), [&](auto& s) {
                return s.second == d.level;
            });
            if (its != by_level.end()) {
                for (size_t k = 0; k < n; ++k) {
                    its->first[k] += d.m[k];
                }
            }
            else {
                by_level.emplace_back(d.m, d.level);
            }
        }
        std::vector<double> comb = by_level.front().first;
        for (size_t l = 1; l < by_level.size(); ++l) {
            for (size_t k = 0; k < n; ++k) {
                const double prod = comb[k] * by_level[l].first[k];
                comb[k]           = comb[k] + by_level[l].first[k] - prod;
            }
        }
        const double tb   = times.back();
        const double diff = std::abs(d.time - tb);
        if (diff <= 1e-15 || diff <= std::max(std::abs(d.time), std::abs(tb)) * 1e-15) {
            std::copy(comb.begin(), comb.end(), mats.end() - long(n));
        }
        else {
            times.push_back(d.time);
            mats.insert(mats.end(), comb.begin(), comb.end());
        }
    }
    times.push_back(std::numeric_limits<double>::max());
    std::vector<double> last(mats.end() - long(n), mats.end());
    mats.insert(mats.end(), last.begin(), last.end());
}

// ------------------------------------------------------------------------------------------ batch
struct RunSpec {
    const mgo_graph_desc* g;
    mgo_sim_cfg cfg;
    double tmax;     // target of this run (ensemble) or the output horizon (graph API)
    bool single;     // mgo_simulate_single: one phase, exact end at tmax
    std::vector<DampIn> extra_damping; // scenario updates added at submission (serving)
};

struct EdgeStructHost {
    std::vector<int32_t> in_ptr, out_ptr, split, seq;
    std::vector<EdgeEnt> in_ent, out_ent;
    std::vector<double> in_min, out_min; // uniform minimum per entry (empty: the graph has no minimum)
    int n = 0, num_edges = 0;
};

// device buffers of one kind of run-time damping lists of a class (see NpiNode), with host copies of what
// a run changes, for reset
struct NpiBufs {
    DevBuf nodes, thr, thr_damp, damp, list, cache_t, cache_m, work, y0;
    std::vector<NpiNode> nodes0;
    std::vector<double> list0, cache_t0, cache_m0;
    void reset(hipStream_t s)
    {
        if (nodes0.empty()) {
            return;
        }
        auto h2d = [&](DevBuf& b, const auto& v, const char* what) {
            hip_check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, s), what);
        };
        h2d(nodes, nodes0, "reset NPI records");
        h2d(list, list0, "reset damping lists");
        h2d(cache_t, cache_t0, "reset damping caches");
        h2d(cache_m, cache_m0, "reset damping caches");
    }
};

struct ClassHost {
    int G, kind;
    std::vector<int> insts; // global instance indices
    int num_nodes = 0;
    // device buffers
    DevBuf y, snap, ypre, rc, yint, lcoef, cbase, cdelta, nodep, dtime, dmat, tnode, dtnode, steps, attempts, out,
        ipool, epool, cvec, cvec_min, mpool, estructs, inst_ids;
    DevBuf fx_rec, fx_pool, fx_dirty;  // fix-up records of the return passes
    DevBuf node_order;                 // node-kernel lane slot -> node (empty: identity)
    // damping lists that change during a run: dynamic NPIs of the contact matrices (nodes with NF_NPI) and
    // out-edge dampings / edge NPIs (nodes with NF_OUT_DAMP, with the damping factor of the last departure)
    NpiBufs npi, oute;
    DevBuf out_d;
    std::vector<NodeP> nodep0; // node records as built (NPIs change their damping-cache counts)
    std::vector<int32_t> order;        // host copy, built by place_nodes
    DevBuf y0;                         // initial state, for reset
    std::vector<double> tnode0, dtnode0; // initial clocks, for reset
    size_t out_doubles = 0;
    DevBuf seg_dev, block_seg, block_first; // launch tables (one segment = the whole class)
    SegDev seg_host{};                       // host copy of the segment descriptor
    LaunchDev launch{};

    // fix-up record versions and dirty marks to "none" (tag -1, below every epoch)
    void clear_fixup_records(hipStream_t s)
    {
        hip_check(hipMemsetAsync(fx_rec.p, 0xff, fx_rec.bytes, s), "clear fix-up records");
        hip_check(hipMemsetAsync(fx_dirty.p, 0xff, fx_dirty.bytes, s), "clear fix-up marks");
    }
};

// blocks of a segment of n nodes with G age groups (BLOCK/WAVE wavefronts of 64/G nodes per block)
int segment_blocks(int n, int G)
{
    const int nodes_per_block = (BLOCK / WAVE) * (WAVE / G);
    return (n + nodes_per_block - 1) / nodes_per_block;
}

// block -> segment table for a list of segments; padded to a multiple of 8 blocks (XCD remap)
void block_tables(const std::vector<int>& seg_nodes, int G, std::vector<int32_t>& block_seg,
                  std::vector<int32_t>& block_first)
{
    block_seg.clear();
    block_first.clear();
    for (size_t k = 0; k < seg_nodes.size(); ++k) {
        block_first.push_back(int32_t(block_seg.size()));
        const int nb = segment_blocks(seg_nodes[k], G);
        block_seg.insert(block_seg.end(), size_t(nb), int32_t(k));
    }
    while (block_seg.empty() || block_seg.size() % 8 != 0) {
        block_seg.push_back(-1);
    }
}

// block -> segment table for segments of different G (mixed launches)
void block_tables_mixed(const std::vector<int>& seg_nodes, const std::vector<int>& seg_G,
                        std::vector<int32_t>& block_seg, std::vector<int32_t>& block_first)
{
    block_seg.clear();
    block_first.clear();
    for (size_t k = 0; k < seg_nodes.size(); ++k) {
        block_first.push_back(int32_t(block_seg.size()));
        const int nb = segment_blocks(seg_nodes[k], seg_G[k]);
        block_seg.insert(block_seg.end(), size_t(nb), int32_t(k));
    }
    while (block_seg.empty() || block_seg.size() % 8 != 0) {
        block_seg.push_back(-1);
    }
}

struct InstHost {
    int cls;
    int node_base;
    int node_stride = 1; // segment node of local node v = node_base + v * node_stride
    int num_nodes;
    int G;
    double t;        // host mirror of the device clock (same arithmetic)
    double dt_graph;
    double tmax;
    double t0;
    int day0, ndays;
    int64_t out_base;
    bool has_edges;
    bool single;
};

int day_floor(double t)
{
    return int(std::floor(t + 1e-10));
}
int day_ceil(double t)
{
    return int(std::ceil(t - 1e-10));
}

static void validate_npis(const mgo_dynamic_npis* dn, int num_nodes, const char* what)
{
    if (!dn) {
        return;
    }
    if (!dn->node_offsets || !dn->node_params) {
        fail(MGO_ERR_INVALID_VALUE, "%s: node_offsets and node_params are required", what);
    }
    if (dn->node_offsets[0] != 0) {
        fail(MGO_ERR_INVALID_VALUE, "%s: node_offsets[0] must be 0", what);
    }
    for (int v = 0; v < num_nodes; ++v) {
        if (dn->node_offsets[v + 1] < dn->node_offsets[v]) {
            fail(MGO_ERR_INVALID_VALUE, "%s: node_offsets must be non-decreasing", what);
        }
        const double* p = dn->node_params + size_t(v) * MGO_NPI_NUM_PARAMS;
        if (!(p[MGO_NPI_DURATION] >= 0) || !(p[MGO_NPI_DELAY] >= 0) || !std::isfinite(p[MGO_NPI_BASE_VALUE]) ||
            !(p[MGO_NPI_DIRECTIVE_BEGIN] < p[MGO_NPI_DIRECTIVE_END])) {
            fail(MGO_ERR_INVALID_VALUE, "%s of node %d: duration and delay must be >= 0, the base value finite and "
                 "the directive begin before its end", what, v);
        }
    }
    const int T = dn->node_offsets[num_nodes];
    if (T > 0 && (!dn->thresholds || !dn->damping_offsets)) {
        fail(MGO_ERR_INVALID_VALUE, "%s: thresholds and damping_offsets are required", what);
    }
    for (int t = 0; t < T; ++t) {
        if (dn->damping_offsets[t + 1] < dn->damping_offsets[t] || dn->damping_offsets[t] < 0) {
            fail(MGO_ERR_INVALID_VALUE, "%s: damping_offsets must be non-decreasing", what);
        }
    }
    if (T > 0 && dn->damping_offsets[T] > dn->damping_offsets[0] &&
        (!dn->damping_values || !dn->damping_levels || !dn->damping_types || !dn->damping_weights)) {
        fail(MGO_ERR_INVALID_VALUE, "%s: damping arrays missing", what);
    }
}

void validate(const mgo_graph_desc* g, const mgo_sim_cfg& cfg)
{
    if (!g) {
        fail(MGO_ERR_INVALID_VALUE, "graph description is NULL");
    }
    if (g->num_nodes < 1) {
        fail(MGO_ERR_INVALID_VALUE, "num_nodes must be >= 1 (got %d)", g->num_nodes);
    }
    if (g->num_groups < 1 || g->num_groups > MGO_MAX_GROUPS) {
        fail(MGO_ERR_UNSUPPORTED, "num_groups must be in 1..%d (got %d)", MGO_MAX_GROUPS, g->num_groups);
    }
    if (!g->node_scalars || !g->age_params || !g->contact_baseline || !g->populations) {
        fail(MGO_ERR_INVALID_VALUE, "node_scalars, age_params, contact_baseline and populations are required");
    }
    if (g->num_edges < 0 || (g->num_edges > 0 && (!g->edge_start || !g->edge_end || !g->edge_coeffs))) {
        fail(MGO_ERR_INVALID_VALUE, "edge arrays missing");
    }
    if (g->damping_offsets && (!g->damping_times || !g->damping_coeffs)) {
        fail(MGO_ERR_INVALID_VALUE, "damping arrays missing");
    }
    validate_npis(g->dynamic_npis, g->num_nodes, "dynamic NPIs");
    validate_npis(g->out_dynamic_npis, g->num_nodes, "edge dynamic NPIs");
    if (g->out_damping_offsets) {
        if (g->out_damping_offsets[0] != 0) {
            fail(MGO_ERR_INVALID_VALUE, "out_damping_offsets[0] must be 0");
        }
        for (int v = 0; v < g->num_nodes; ++v) {
            if (g->out_damping_offsets[v + 1] < g->out_damping_offsets[v]) {
                fail(MGO_ERR_INVALID_VALUE, "out_damping_offsets must be non-decreasing");
            }
        }
        if (g->out_damping_offsets[g->num_nodes] > 0 && (!g->out_damping_times || !g->out_damping_coeffs)) {
            fail(MGO_ERR_INVALID_VALUE, "out-edge damping arrays missing");
        }
    }
    const int k = cfg.integrator.kind;
    if (k < MGO_INTEGRATOR_CASH_KARP54 || k > MGO_INTEGRATOR_EULER) {
        fail(MGO_ERR_INVALID_VALUE, "unknown integrator kind %d", k);
    }
    if ((k == MGO_INTEGRATOR_CASH_KARP54 || k == MGO_INTEGRATOR_RKF45) &&
        (!(cfg.integrator.dt_min >= 0) || !(cfg.integrator.dt_min <= cfg.integrator.dt_max) ||
         !(cfg.integrator.abs_tol >= 0) || !(cfg.integrator.rel_tol >= 0))) {
        fail(MGO_ERR_INVALID_VALUE, "integrator bounds/tolerances invalid");
    }
    if (!(cfg.dt_node > 0) || !std::isfinite(cfg.t0)) {
        fail(MGO_ERR_INVALID_VALUE, "dt_node must be > 0 and t0 finite");
    }
    const int G = g->num_groups;
    for (int e = 0; e < g->num_edges; ++e) {
        if (g->edge_start[e] < 0 || g->edge_start[e] >= g->num_nodes || g->edge_end[e] < 0 ||
            g->edge_end[e] >= g->num_nodes) {
            fail(MGO_ERR_OUT_OF_RANGE, "edge %d references node out of range", e);
        }
    }
    for (int v = 0; v < g->num_nodes; ++v) {
        const double* a = g->age_params + size_t(v) * MGO_NUM_AGE_PARAMS * G;
        for (int q = 0; q < 5 * G; ++q) {
            if (!(a[q] > 0)) {
                fail(MGO_ERR_INVALID_VALUE, "node %d: stay times must be > 0", v);
            }
        }
        for (int q = 0; q < G * NC; ++q) {
            if (!std::isfinite(g->populations[size_t(v) * G * NC + q])) {
                fail(MGO_ERR_INVALID_VALUE, "node %d: populations must be finite", v);
            }
        }
    }
}

// lane coefficients: the reference's subexpressions (cpp/models/ode_secir/model.h:93-212)
void lane_coefficients(const mgo_graph_desc* g, int v, int a, double* c)
{
    const int G       = g->num_groups;
    auto P            = [&](int k) {
        return g->age_params[(size_t(v) * MGO_NUM_AGE_PARAMS + k) * G + a];
    };
    const double TE   = P(MGO_A_TIME_EXPOSED), TINS = P(MGO_A_TIME_INFECTED_NO_SYMPTOMS),
                 TISY = P(MGO_A_TIME_INFECTED_SYMPTOMS), TISEV = P(MGO_A_TIME_INFECTED_SEVERE),
                 TICR = P(MGO_A_TIME_INFECTED_CRITICAL);
    const double Rr = P(MGO_A_RECOVERED_PER_INFECTED_NO_SYMPTOMS), sev = P(MGO_A_SEVERE_PER_INFECTED_SYMPTOMS),
                 cps = P(MGO_A_CRITICAL_PER_SEVERE), dps = P(MGO_A_DEATHS_PER_SEVERE),
                 dpc = P(MGO_A_DEATHS_PER_CRITICAL);
    c[LC_E]        = (1.0 / TE);
    c[LC_INS_ISY]  = (1 - Rr) * (1 / TINS);
    c[LC_INS_R]    = Rr * (1.0 / TINS);
    c[LC_ISY_ISEV] = sev / TISY;
    c[LC_ISY_R]    = (1 - sev) / TISY;
    c[LC_ISEV_ICR] = cps / TISEV;
    c[LC_ISEV_R]   = (1.0 - cps - dps) / TISEV;
    c[LC_ISEV_D]   = (dps + (cps - cps)) / TISEV;
    c[LC_ICR_D]    = dpc / TICR;
    c[LC_ICR_R]    = (1.0 - dpc) / TICR;
    c[LC_TP]       = P(MGO_A_TRANSMISSION_PROBABILITY_ON_CONTACT);
    c[LC_REL]      = P(MGO_A_RELATIVE_TRANSMISSION_NO_SYMPTOMS);
    c[LC_RISK]     = P(MGO_A_RISK_OF_INFECTION_FROM_SYMPTOMATIC);
    c[LC_MAXRISK]  = P(MGO_A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC);
    c[LC_TTR]      = (1 - Rr) / TINS;
    c[LC_CPS]      = cps;
    c[LC_DPS]      = dps;
    c[LC_TISEV]    = TISEV;
}

// CSR of one graph: Graph::add_edge ordering (sorted by (start, end), duplicates replaced by the later
// edge, graph.h:231-241); edges without a positive coefficient never move anyone and are dropped.
void build_edges(const mgo_graph_desc* g, EdgeStructHost& es, std::vector<double>& cvec, std::vector<double>& cvec_min)
{
    const int n  = g->num_nodes;
    const int G  = g->num_groups;
    const int nv = G * NC;
    // (start, end) keys sorted; a repeated key keeps its last occurrence (Graph::add_edge replaces)
    std::vector<std::pair<uint64_t, int>> keys(size_t(std::max(g->num_edges, 0)));
    for (int e = 0; e < g->num_edges; ++e) {
        keys[size_t(e)] = {(uint64_t(uint32_t(g->edge_start[e])) << 32) | uint32_t(g->edge_end[e]), e};
    }
    std::sort(keys.begin(), keys.end());
    struct E {
        int s, d;
        double cval, cmin;
        int cidx;
    };
    std::vector<E> edges;
    edges.reserve(keys.size());
    const int len = (g->edge_coeff_len == 1) ? 1 : nv;
    for (size_t k = 0; k < keys.size(); ++k) {
        if (k + 1 < keys.size() && keys[k + 1].first == keys[k].first) {
            continue; // a later occurrence of the same (start, end) follows
        }
        const int src   = int(keys[k].first >> 32);
        const int dst   = int(uint32_t(keys[k].first));
        const double* c  = g->edge_coeffs + size_t(keys[k].second) * len;
        const double* mn = g->edge_coeff_minimum ? g->edge_coeff_minimum + size_t(keys[k].second) * len : nullptr;
        bool any_pos = false, uniform = true;
        for (int k = 0; k < len; ++k) {
            // an edge whose coefficients can never be positive (B - D (B - M) with B, M <= 0) never departs
            any_pos = any_pos || (c[k] > 0.0) || (mn && mn[k] > 0.0);
            uniform = uniform && (c[k] == c[0]) && (!mn || mn[k] == mn[0]);
        }
        if (!any_pos) {
            continue;
        }
        E ed{src, dst, c[0], mn ? mn[0] : 0.0, -1};
        if (!uniform) {
            ed.cidx = int(cvec.size() / size_t(nv));
            cvec.insert(cvec.end(), c, c + nv);
            if (mn) {
                cvec_min.insert(cvec_min.end(), mn, mn + nv);
            }
            else {
                cvec_min.insert(cvec_min.end(), size_t(nv), 0.0);
            }
        }
        edges.push_back(ed);
    }
    es.n         = n;
    es.num_edges = int(edges.size());
    es.out_ptr.assign(size_t(n) + 1, 0);
    es.in_ptr.assign(size_t(n) + 1, 0);
    for (auto& e : edges) {
        es.out_ptr[size_t(e.s) + 1]++;
        es.in_ptr[size_t(e.d) + 1]++;
    }
    for (int v = 0; v < n; ++v) {
        es.out_ptr[size_t(v) + 1] += es.out_ptr[size_t(v)];
        es.in_ptr[size_t(v) + 1] += es.in_ptr[size_t(v)];
    }
    es.out_ent.resize(edges.size());
    es.in_ent.resize(edges.size());
    if (g->edge_coeff_minimum) {
        es.out_min.assign(edges.size(), 0.0);
        es.in_min.assign(edges.size(), 0.0);
    }
    es.split.assign(size_t(n), 0);
    es.seq.clear();
    std::vector<int> ofill(es.out_ptr.begin(), es.out_ptr.end() - 1), ifill(es.in_ptr.begin(), es.in_ptr.end() - 1);
    // edges are sorted by (s, d): out lists come out sorted by d, in lists sorted by s
    for (auto& e : edges) {
        const int ko    = ofill[size_t(e.s)]++;
        const int ki    = ifill[size_t(e.d)]++;
        es.out_ent[size_t(ko)] = EdgeEnt{e.cval, e.d, e.cidx, ki, 0};
        es.in_ent[size_t(ki)]  = EdgeEnt{e.cval, e.s, e.cidx, ko, 0};
        if (!es.out_min.empty()) {
            es.out_min[size_t(ko)] = e.cmin;
            es.in_min[size_t(ki)]  = e.cmin;
        }
        if (e.s < e.d) {
            es.split[size_t(e.d)]++;
        }
        es.seq.push_back(e.s);
        es.seq.push_back(e.d);
        es.seq.push_back(ko);
    }
}

// ------------------------------------------------------------------------------------------ dynamic NPIs
// Host build of the per-node NPI data of one class (see NpiNode and mgo_npi.h).
struct NpiBuild {
    struct Node {
        size_t node;
        NpiNode rec{};
        std::vector<double> list, cache_t, cache_m;
    };
    std::vector<Node> nodes;
    std::vector<double> thr, damp;  // threshold values (per node descending); dampings [3 + G]
    std::vector<int32_t> thr_damp;  // first damping of each threshold (+ a final end marker)
};

// One damping list with its NPIs: node v's contact matrix (contact: n = G*G, NPIs g->dynamic_npis) or the
// node's out-edge coefficients (n = G per-group values, NPIs g->out_dynamic_npis; entries of dl are G long).
static void add_npi_node(NpiBuild& b, size_t node, const mgo_graph_desc* g, int v, int G, bool contact,
                         const std::vector<DampIn>& dl)
{
    const mgo_dynamic_npis* dn = contact ? g->dynamic_npis : g->out_dynamic_npis;
    const int n                = contact ? G * G : G;
    NpiBuild::Node nd;
    nd.node    = node;
    NpiNode& r = nd.rec;
    r.duration = 14.0;
    r.delay    = 0.0;
    r.base     = 1.0;
    r.begin    = -DBL_MAX;
    r.end      = DBL_MAX;
    if (dn) {
        const double* par = dn->node_params + size_t(v) * MGO_NPI_NUM_PARAMS;
        r.duration        = par[MGO_NPI_DURATION];
        r.delay           = par[MGO_NPI_DELAY];
        r.base            = par[MGO_NPI_BASE_VALUE];
        r.begin           = par[MGO_NPI_DIRECTIVE_BEGIN];
        r.end             = par[MGO_NPI_DIRECTIVE_END];
    }
    r.value      = -DBL_MAX;
    r.t_end      = 0.0;
    double total = 0.0; // Populations::get_total of the model (summed in storage order)
    for (int k = 0; k < G * NC; ++k) {
        total += g->populations[size_t(v) * G * NC + size_t(k)];
    }
    r.total = total;
    // DynamicNPIs::set_threshold: sorted by value, descending; an equal value replaces the earlier one
    std::vector<std::pair<double, int>> thr;
    if (dn) {
        for (int t = dn->node_offsets[v]; t < dn->node_offsets[v + 1]; ++t) {
            const double val = dn->thresholds[t];
            auto it          = thr.begin();
            while (it != thr.end() && it->first > val) {
                ++it;
            }
            if (it != thr.end() && !(val > it->first)) {
                it->second = t;
            }
            else {
                thr.insert(it, std::make_pair(val, t));
            }
        }
    }
    r.thr_first = int32_t(b.thr.size());
    r.thr_count = int32_t(thr.size());
    for (auto& tv : thr) {
        b.thr.push_back(tv.first);
        b.thr_damp.push_back(int32_t(b.damp.size() / size_t(3 + G)));
        for (int d = dn->damping_offsets[tv.second]; d < dn->damping_offsets[tv.second + 1]; ++d) {
            b.damp.push_back(dn->damping_values[d]);
            b.damp.push_back(double(dn->damping_levels[d]));
            b.damp.push_back(double(dn->damping_types[d]));
            b.damp.insert(b.damp.end(), dn->damping_weights + size_t(d) * G, dn->damping_weights + size_t(d + 1) * G);
        }
    }
    // damping list: the initial dampings in add order, room for MGO_NPI_LIST_EXTRA more
    r.list_cap = int32_t(dl.size()) + MGO_NPI_LIST_EXTRA;
    nd.list.assign(size_t(r.list_cap) * size_t(npi_entry_size(n)), 0.0);
    int count = 0;
    for (const DampIn& d : dl) {
        npi_list_add(nd.list.data(), count, r.list_cap, n, d.time, d.level, d.type, d.m.data());
    }
    r.list_count = count;
    nd.cache_t.assign(size_t(r.list_cap) + 2, 0.0);
    nd.cache_m.assign((size_t(r.list_cap) + 2) * size_t(n), 0.0);
    std::vector<double> work(size_t(npi_work_size(n, r.list_cap)));
    const int m = npi_build_cache(nd.list.data(), count, n, r.list_cap, nd.cache_t.data(), nd.cache_m.data(),
                                  work.data());
    if (m < 0) {
        fail(MGO_ERR_UNSUPPORTED, "node %d: more than %d damping levels", v, NPI_LEVELS);
    }
    r.cache_count = m;
    b.nodes.push_back(std::move(nd));
}

struct Batch {
    hipStream_t stream = nullptr;
    std::deque<ClassHost> classes;
    std::vector<InstHost> inst;
    DevBuf inst_dev;
    DevBuf cyc; // [CYC_FLAGS] cycle flags (clock kernel, fixup_check), shared by the classes' launches
    int64_t cycles_run = 0;
    double last_ms     = 0;
    hipGraphExec_t graph_exec = nullptr;
    bool use_graph            = true;
    bool concurrent_classes   = true; // classes of a cycle on their own streams (off: one stream)
    bool profile              = false;
    bool sequential_fixup     = false;
    size_t device_bytes       = 0;
    std::vector<InstDev> inst0; // instance table at creation (tmax set), for reset
    std::deque<std::vector<double>> staging; // host sources of uploads in flight during build
    // kernel timing of the last run (profile mode)
    double node_ms = 0, edge_ms = 0;
    int node_launches = 0, edge_launches = 0;

    // classes of a cycle run concurrently on these streams (one per class, default priority), forked from
    // and joined back into `stream` (also inside the captured cycle graph)
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> side_done;
    hipEvent_t fork_ev = nullptr;

    ~Batch()
    {
        if (graph_exec) {
            (void)hipGraphExecDestroy(graph_exec);
        }
        for (auto e : side_done) {
            (void)hipEventDestroy(e);
        }
        for (auto st : side) {
            (void)hipStreamDestroy(st);
        }
        if (fork_ev) {
            (void)hipEventDestroy(fork_ev);
        }
    }

    void make_side_streams(size_t n)
    {
        if (!fork_ev) {
            hip_check(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming), "fork event");
        }
        while (side.size() < n) {
            hipStream_t st = nullptr;
            hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "class stream");
            side.push_back(st);
            hipEvent_t e = nullptr;
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "class event");
            side_done.push_back(e);
        }
    }

    size_t class_index(std::map<std::pair<int, int>, int>& cmap, int G, int kind)
    {
        auto it = cmap.find({G, kind});
        if (it != cmap.end()) {
            return size_t(it->second);
        }
        cmap[{G, kind}] = int(classes.size());
        classes.emplace_back();
        classes.back().G    = G;
        classes.back().kind = kind;
        return classes.size() - 1;
    }

    int inst_offset = 0; // global index of instance 0 (serving: the request's slot in the shared table)

    int tile = 1; // runs per node-major tile (see Tiling in mgo_internal.h)

    void build(std::vector<RunSpec>& runs, hipStream_t s, int tile_runs = 1)
    {
        stream = s;
        tile   = std::max(1, tile_runs);
        for (auto& r : runs) {
            validate(r.g, r.cfg);
            if (!r.single && !(r.cfg.dt_graph > 0)) {
                fail(MGO_ERR_INVALID_VALUE, "dt_graph must be > 0");
            }
        }
        std::map<std::pair<int, int>, int> cmap;
        inst.resize(runs.size());
        for (size_t i = 0; i < runs.size(); ++i) {
            const auto& r = runs[i];
            const int ci  = int(class_index(cmap, r.g->num_groups, r.cfg.integrator.kind));
            auto& C       = classes[size_t(ci)];
            InstHost& I   = inst[i];
            I.cls         = ci;
            I.G           = r.g->num_groups;
            I.num_nodes   = r.g->num_nodes;
            C.insts.push_back(int(i));
            I.t0       = r.cfg.t0;
            I.t        = r.cfg.t0;
            I.dt_graph = r.single ? INFINITY : r.cfg.dt_graph;
            I.tmax     = r.cfg.t0;
            I.single   = r.single;
            I.day0     = day_ceil(r.cfg.t0);
            I.ndays    = std::max(0, day_floor(r.tmax) - I.day0 + 1);
        }
        for (auto& C : classes) {
            place_nodes(C, runs);
        }
        // sources of the asynchronous uploads below; they outlive the handler
        const std::vector<int32_t> cyc_init(CYC_FLAGS, 0);
        std::vector<InstDev> idev(runs.size());
        try {
            cyc.upload(cyc_init, stream);
            for (size_t ci = 0; ci < classes.size(); ++ci) {
                build_class(ci, runs, idev);
            }
            inst_dev.upload(idev, stream);
        }
        catch (...) {
            // an allocation failed part-way (a full serving arena): the uploads already queued must land
            // before the caller frees this batch, or they would write into memory handed to someone else
            (void)hipStreamSynchronize(stream);
            throw;
        }
        device_bytes += inst_dev.bytes;
        hip_check(hipStreamSynchronize(stream), "batch upload");
        staging.clear();
    }

    // Segment node indices of a class: runs that share a graph (same edge arrays and node count) are
    // interleaved node-major in tiles of up to `tile` runs; the others get contiguous ranges.
    void place_nodes(ClassHost& C, const std::vector<RunSpec>& runs)
    {
        std::map<std::tuple<const void*, const void*, const void*, int, int, bool>, std::vector<int>> groups;
        std::vector<std::vector<int>*> order;
        for (int ii : C.insts) {
            const auto* g = runs[size_t(ii)].g;
            auto key = std::make_tuple((const void*)g->edge_start, (const void*)g->edge_end, (const void*)g->edge_coeffs,
                                       g->num_edges, g->num_nodes, runs[size_t(ii)].single);
            auto it = groups.find(key);
            if (it == groups.end()) {
                it = groups.emplace(key, std::vector<int>{}).first;
                order.push_back(&it->second);
            }
            it->second.push_back(ii);
        }
        bool shared = false;
        for (auto* grp : order) {
            shared = shared || grp->size() > 1;
            for (size_t first = 0; first < grp->size(); first += size_t(tile)) {
                const int k = int(std::min(grp->size() - first, size_t(tile)));
                const int n = inst[size_t((*grp)[first])].num_nodes;
                for (int j = 0; j < k; ++j) {
                    InstHost& I   = inst[size_t((*grp)[first + size_t(j)])];
                    I.node_base   = C.num_nodes + j;
                    I.node_stride = k;
                }
                C.num_nodes += n * k;
            }
        }
        // node-kernel order: each graph node of a group of runs sharing a graph over all its runs
        C.order.clear();
        if (shared) {
            for (auto* grp : order) {
                const int n = inst[size_t((*grp)[0])].num_nodes;
                for (int v = 0; v < n; ++v) {
                    for (int ii : *grp) {
                        const InstHost& I = inst[size_t(ii)];
                        C.order.push_back(int32_t(I.node_base + v * I.node_stride));
                    }
                }
            }
        }
    }

    // device buffers of one kind of run-time damping lists (n coefficients per entry); for the contact kind
    // the nodes' damping caches point into them
    void upload_npis(ClassHost& C, NpiBufs& B, NpiBuild& b, int n, bool contact, std::vector<NodeP>& nodep)
    {
        B.nodes0.clear();
        if (b.nodes.empty()) {
            return;
        }
        const int G = C.G;
        b.thr_damp.push_back(int32_t(b.damp.size() / size_t(3 + G)));
        size_t nl = 0, nct = 0, nw = 0;
        for (auto& nd : b.nodes) {
            nl += nd.list.size();
            nct += nd.cache_t.size();
            nw += size_t(npi_work_size(n, nd.rec.list_cap));
        }
        B.list0.clear();
        B.cache_t0.clear();
        B.cache_m0.clear();
        B.list0.reserve(nl);
        B.cache_t0.reserve(nct);
        B.cache_m0.reserve(nct * size_t(n));
        for (auto& nd : b.nodes) {
            B.list0.insert(B.list0.end(), nd.list.begin(), nd.list.end());
            B.cache_t0.insert(B.cache_t0.end(), nd.cache_t.begin(), nd.cache_t.end());
            B.cache_m0.insert(B.cache_m0.end(), nd.cache_m.begin(), nd.cache_m.end());
        }
        B.list.upload(B.list0, stream);
        B.cache_t.upload(B.cache_t0, stream);
        B.cache_m.upload(B.cache_m0, stream);
        B.work.alloc(nw * sizeof(double));
        if (contact) {
            B.y0.alloc(b.nodes.size() * size_t(G) * NC * sizeof(double));
        }
        B.thr.upload(b.thr, stream);
        B.thr_damp.upload(b.thr_damp, stream);
        B.damp.upload(b.damp, stream);
        B.nodes0.assign(size_t(C.num_nodes), NpiNode{});
        size_t ol = 0, oc = 0, ow = 0;
        for (size_t k = 0; k < b.nodes.size(); ++k) {
            NpiBuild::Node& nd = b.nodes[k];
            NpiNode r          = nd.rec;
            r.list             = B.list.as<double>() + ol;
            r.cache_t          = B.cache_t.as<double>() + oc;
            r.cache_m          = B.cache_m.as<double>() + oc * size_t(n);
            r.work             = B.work.as<double>() + ow;
            r.y0               = contact ? B.y0.as<double>() + k * size_t(G) * NC : nullptr;
            ol += nd.list.size();
            oc += nd.cache_t.size();
            ow += size_t(npi_work_size(n, r.list_cap));
            B.nodes0[nd.node] = r;
            if (contact) {
                NodeP& np    = nodep[nd.node];
                np.damp_time = r.cache_t;
                np.damp_mat  = r.cache_m;
                np.damp_cnt  = r.cache_count;
            }
        }
        B.nodes.upload(B.nodes0, stream);
        device_bytes += B.list.bytes + B.cache_t.bytes + B.cache_m.bytes + B.work.bytes + B.y0.bytes + B.thr.bytes +
                        B.thr_damp.bytes + B.damp.bytes + B.nodes.bytes;
    }

    void build_class(size_t ci, std::vector<RunSpec>& runs, std::vector<InstDev>& idev)
    {
        ClassHost& C = classes[ci];
        const int G  = C.G;
        const size_t N = size_t(C.num_nodes);
        std::vector<double> y(N * G * NC), lcoef(N * G * NUM_LC), cbase(N * G * G), cdelta(N * G * G), tnode(N),
            dtnode(N), dtime, dmat;
        std::vector<NodeP> nodep(N);
        std::vector<int32_t> ipool, inst_ids;
        std::vector<EdgeEnt> epool;
        std::vector<double> cvec, cvec_min, mpool;
        struct StructOff {
            size_t in_ptr, out_ptr, split, seq, in_ent, out_ent;
            int64_t in_min = -1, out_min = -1; // offsets into mpool, -1 = no minimum
            int n, num_edges;
        };
        std::vector<StructOff> soffs;
        std::vector<size_t> damp_off(N, 0);
        NpiBuild npi, outd;
        std::map<std::tuple<const void*, const void*, const void*, const void*, int, int>, int> struct_cache;
        int64_t out_total = 0;
        int64_t fx_total  = 0;
        int64_t pool_slots = 0;
        for (int ii : C.insts) {
            RunSpec& r    = runs[size_t(ii)];
            InstHost& I   = inst[size_t(ii)];
            const auto* g = r.g;
            inst_ids.push_back(ii);
            for (int v = 0; v < g->num_nodes; ++v) {
                const size_t node = size_t(I.node_base) + size_t(v) * size_t(I.node_stride);
                const double* sc  = g->node_scalars + size_t(v) * MGO_NUM_SCALARS;
                NodeP& np         = nodep[node];
                np.start_day      = sc[MGO_P_START_DAY];
                np.seasonality    = sc[MGO_P_SEASONALITY];
                np.ttc            = sc[MGO_P_TEST_AND_TRACE_CAPACITY];
                np.ttc_hi         = sc[MGO_P_TEST_AND_TRACE_CAPACITY] * sc[MGO_P_TEST_AND_TRACE_MAX_RISK];
                np.icu            = sc[MGO_P_ICU_CAPACITY];
                np.nondetect      = sc[MGO_P_COMMUTER_NONDETECTION];
                np.det_start      = sc[MGO_P_START_COMMUTER_DETECTION];
                np.det_end        = sc[MGO_P_END_COMMUTER_DETECTION];
                np.inst           = inst_offset + ii;
                np.local          = v;
                np.flags          = 0;
                if (np.ttc < DBL_MAX) {
                    np.flags |= NF_TTC_ACTIVE;
                }
                if (np.icu < DBL_MAX) {
                    np.flags |= NF_ICU_ACTIVE;
                }
                if (np.det_start < np.det_end) {
                    np.flags |= NF_DETECTION;
                }
                for (int a = 0; a < G; ++a) {
                    lane_coefficients(g, v, a, &lcoef[(node * G + size_t(a)) * NUM_LC]);
                    for (int j = 0; j < G; ++j) {
                        const size_t k = (size_t(v) * G + size_t(a)) * G + size_t(j);
                        const double b = g->contact_baseline[k];
                        const double m = g->contact_minimum ? g->contact_minimum[k] : 0.0;
                        cbase[(node * G + size_t(a)) * G + size_t(j)]  = b;
                        cdelta[(node * G + size_t(a)) * G + size_t(j)] = b - m;
                    }
                }
                std::copy(g->populations + size_t(v) * G * NC, g->populations + size_t(v + 1) * G * NC,
                          y.begin() + long(node * G * NC));
                tnode[node]  = r.cfg.t0;
                dtnode[node] = r.cfg.dt_node;
                // damping cache
                std::vector<DampIn> dl;
                if (g->damping_offsets) {
                    for (int d = g->damping_offsets[v]; d < g->damping_offsets[v + 1]; ++d) {
                        DampIn di;
                        di.time  = g->damping_times[d];
                        di.level = g->damping_levels ? g->damping_levels[d] : 0;
                        di.type  = g->damping_types ? g->damping_types[d] : 0;
                        di.m.assign(g->damping_coeffs + size_t(d) * G * G, g->damping_coeffs + size_t(d + 1) * G * G);
                        dl.push_back(di);
                    }
                }
                for (auto& d : r.extra_damping) {
                    dl.push_back(d);
                }
                damp_off[node] = dtime.size();
                np.damp_cnt    = 0;
                np.damp_time   = nullptr;
                np.damp_mat    = nullptr;
                // out-edge coefficient dampings and edge NPIs (shared by the node's out-edges)
                {
                    std::vector<DampIn> ol;
                    if (g->out_damping_offsets) {
                        for (int d = g->out_damping_offsets[v]; d < g->out_damping_offsets[v + 1]; ++d) {
                            DampIn di;
                            di.time  = g->out_damping_times[d];
                            di.level = g->out_damping_levels ? g->out_damping_levels[d] : 0;
                            di.type  = g->out_damping_types ? g->out_damping_types[d] : 0;
                            di.m.assign(g->out_damping_coeffs + size_t(d) * G, g->out_damping_coeffs + size_t(d + 1) * G);
                            ol.push_back(di);
                        }
                    }
                    const mgo_dynamic_npis* on = g->out_dynamic_npis;
                    if (!r.single && (!ol.empty() || (on && on->node_offsets[v + 1] > on->node_offsets[v]))) {
                        add_npi_node(outd, node, g, v, G, false, ol);
                        np.flags |= NF_OUT_DAMP;
                    }
                }
                const mgo_dynamic_npis* dn = g->dynamic_npis;
                if (dn && dn->node_offsets[v + 1] > dn->node_offsets[v]) {
                    add_npi_node(npi, node, g, v, G, true, dl);
                    np.flags |= NF_NPI | NF_DAMPED;
                }
                else if (!dl.empty()) {
                    std::vector<double> tt, mm;
                    damping_cache(dl, G, tt, mm);
                    np.flags |= NF_DAMPED;
                    np.damp_cnt = int(tt.size());
                    dtime.insert(dtime.end(), tt.begin(), tt.end());
                    dmat.insert(dmat.end(), mm.begin(), mm.end());
                }
            }
            // edge structure (shared between runs that pass the same edge arrays)
            InstDev& D = idev[size_t(ii)];
            D.estruct  = -1;
            I.has_edges = false;
            if (!r.single && g->num_edges > 0) {
                auto key = std::make_tuple((const void*)g->edge_start, (const void*)g->edge_end,
                                           (const void*)g->edge_coeffs, (const void*)g->edge_coeff_minimum,
                                           g->num_edges, g->num_nodes);
                auto it  = struct_cache.find(key);
                int sidx;
                if (it != struct_cache.end()) {
                    sidx = it->second;
                }
                else {
                    EdgeStructHost es;
                    build_edges(g, es, cvec, cvec_min);
                    StructOff so{};
                    so.n         = es.n;
                    so.num_edges = es.num_edges;
                    so.in_ptr    = ipool.size();
                    ipool.insert(ipool.end(), es.in_ptr.begin(), es.in_ptr.end());
                    so.out_ptr = ipool.size();
                    ipool.insert(ipool.end(), es.out_ptr.begin(), es.out_ptr.end());
                    so.split = ipool.size();
                    ipool.insert(ipool.end(), es.split.begin(), es.split.end());
                    so.seq = ipool.size();
                    ipool.insert(ipool.end(), es.seq.begin(), es.seq.end());
                    so.in_ent = epool.size();
                    epool.insert(epool.end(), es.in_ent.begin(), es.in_ent.end());
                    so.out_ent = epool.size();
                    epool.insert(epool.end(), es.out_ent.begin(), es.out_ent.end());
                    if (!es.in_min.empty()) {
                        so.in_min = int64_t(mpool.size());
                        mpool.insert(mpool.end(), es.in_min.begin(), es.in_min.end());
                        so.out_min = int64_t(mpool.size());
                        mpool.insert(mpool.end(), es.out_min.begin(), es.out_min.end());
                    }
                    sidx = int(soffs.size());
                    soffs.push_back(so);
                    struct_cache[key] = sidx;
                }
                if (soffs[size_t(sidx)].num_edges > 0) {
                    D.estruct   = sidx;
                    I.has_edges = true;
                }
            }
            D.fx_base      = fx_total;
            D.fx_pool_base = int64_t(pool_slots) * FX_SLOTS * G * NC;
            if (D.estruct >= 0) {
                fx_total += soffs[size_t(D.estruct)].num_edges;
                ++pool_slots;
            }
            D.t         = r.cfg.t0;
            D.dt_graph  = I.dt_graph;
            D.tmax      = r.cfg.t0;
            D.abs_tol   = r.cfg.integrator.abs_tol;
            D.rel_tol   = r.cfg.integrator.rel_tol;
            D.dt_min    = r.cfg.integrator.dt_min;
            D.dt_max    = r.cfg.integrator.dt_max;
            D.node_base = I.node_base;
            D.node_stride = I.node_stride;
            D.num_nodes = I.num_nodes;
            D.day0      = I.day0;
            D.ndays     = I.ndays;
            D.exact_end = r.single ? 1 : 0;
            D.fx_sequential = sequential_fixup ? 1 : 0;
            D.seg           = 0;
            I.out_base  = out_total;
            out_total += int64_t(I.ndays) * I.num_nodes * G * NC;
        }
        // initial day: interpolate_simulation_result keeps t0 if t0 >= day0 (analyze_result.h:228-233)
        std::vector<double> out0; // written on the device below
        C.out_doubles = size_t(out_total);
        auto up = [&](DevBuf& b, const auto& v) {
            b.upload(v, stream);
            device_bytes += b.bytes;
        };
        up(C.y, y);
        up(C.y0, y);
        C.tnode0  = tnode;
        C.dtnode0 = dtnode;
        C.snap.alloc(N * G * SNAP_STRIDE * sizeof(double));
        C.ypre.alloc(y.size() * sizeof(double));
        C.rc.alloc(N * G * RC_STRIDE * sizeof(double));
        C.yint.alloc(N * G * YINT_STRIDE * sizeof(double));
        hip_check(hipMemsetAsync(C.yint.p, 0xff, C.yint.bytes, stream), "clear yint"); // NaN weights: no step yet
        device_bytes += C.snap.bytes + C.ypre.bytes + C.rc.bytes + C.yint.bytes;
        up(C.lcoef, lcoef);
        up(C.cbase, cbase);
        up(C.cdelta, cdelta);
        up(C.dtime, dtime);
        up(C.dmat, dmat);
        for (size_t node = 0; node < N; ++node) {
            if (nodep[node].damp_cnt > 0 && !(nodep[node].flags & NF_NPI)) {
                nodep[node].damp_time = C.dtime.as<double>() + damp_off[node];
                nodep[node].damp_mat  = C.dmat.as<double>() + damp_off[node] * G * G;
            }
        }
        upload_npis(C, C.npi, npi, G * G, true, nodep);
        upload_npis(C, C.oute, outd, G, false, nodep);
        if (!C.oute.nodes0.empty()) {
            C.out_d.alloc(N * size_t(G) * sizeof(double));
            hip_check(hipMemsetAsync(C.out_d.p, 0, C.out_d.bytes, stream), "damping factors");
            device_bytes += C.out_d.bytes;
        }
        C.nodep0 = nodep;
        up(C.nodep, nodep);
        up(C.tnode, tnode);
        up(C.dtnode, dtnode);
        std::vector<uint32_t> zeros(N, 0);
        up(C.steps, zeros);
        up(C.attempts, zeros);
        C.out.alloc(C.out_doubles * sizeof(double));
        device_bytes += C.out.bytes;
        hip_check(hipMemsetAsync(C.out.p, 0, C.out.bytes, stream), "memset out");
        up(C.ipool, ipool);
        up(C.epool, epool);
        up(C.cvec, cvec);
        up(C.cvec_min, cvec_min);
        up(C.mpool, mpool);
        std::vector<EdgeStructDev> estructs(soffs.size());
        for (size_t k = 0; k < soffs.size(); ++k) {
            const auto& so          = soffs[k];
            EdgeStructDev& sd       = estructs[k];
            const int32_t* ib       = C.ipool.as<int32_t>();
            const EdgeEnt* eb       = C.epool.as<EdgeEnt>();
            sd.in_ptr               = ib + so.in_ptr;
            sd.out_ptr              = ib + so.out_ptr;
            sd.split                = ib + so.split;
            sd.seq                  = ib + so.seq;
            sd.in_ent               = eb + so.in_ent;
            sd.out_ent              = eb + so.out_ent;
            sd.cvec                 = C.cvec.as<double>();
            sd.cvec_min             = C.cvec_min.as<double>();
            sd.in_min               = so.in_min >= 0 ? C.mpool.as<double>() + so.in_min : nullptr;
            sd.out_min              = so.out_min >= 0 ? C.mpool.as<double>() + so.out_min : nullptr;
            sd.n                    = so.n;
            sd.num_edges            = so.num_edges;
        }
        up(C.estructs, estructs);
        up(C.inst_ids, inst_ids);
        C.fx_rec.alloc(size_t(std::max<int64_t>(fx_total, 1)) * 2 * sizeof(FxRec));
        C.fx_pool.alloc(size_t(std::max<int64_t>(pool_slots, 1)) * FX_SLOTS * G * NC * sizeof(double));
        C.fx_dirty.alloc(2 * std::max<size_t>(N, 1) * sizeof(int32_t));
        device_bytes += C.fx_rec.bytes + C.fx_pool.bytes + C.fx_dirty.bytes;
        C.clear_fixup_records(stream);
        for (int ii : C.insts) {
            idev[size_t(ii)].out = C.out.as<double>() + inst[size_t(ii)].out_base;
        }
        // day-0 values
        for (int ii : C.insts) {
            InstHost& I = inst[size_t(ii)];
            if (I.ndays > 0 && I.t0 >= double(I.day0) - 1e-10 && I.t0 <= double(I.day0)) {
                const size_t nd = size_t(I.num_nodes) * G * NC;
                if (I.node_stride == 1) {
                    hip_check(hipMemcpyAsync(C.out.as<double>() + I.out_base,
                                             C.y.as<double>() + size_t(I.node_base) * G * NC, nd * sizeof(double),
                                             hipMemcpyDeviceToDevice, stream),
                              "day0 copy");
                }
                else {
                    staging.emplace_back();
                    std::vector<double>& d0 = staging.back();
                    d0.resize(nd);
                    for (int v = 0; v < I.num_nodes; ++v) {
                        const size_t node = size_t(I.node_base) + size_t(v) * size_t(I.node_stride);
                        std::copy(y.begin() + long(node * G * NC), y.begin() + long((node + 1) * G * NC),
                                  d0.begin() + long(size_t(v) * G * NC));
                    }
                    hip_check(hipMemcpyAsync(C.out.as<double>() + I.out_base, d0.data(), nd * sizeof(double),
                                             hipMemcpyHostToDevice, stream),
                              "day0 copy");
                }
            }
        }
        SegDev d{};
        d.G         = G;
        d.num_nodes = C.num_nodes;
        d.y         = C.y.as<double>();
        d.snap      = C.snap.as<double>();
        d.ypre      = C.ypre.as<double>();
        d.rc        = C.rc.as<double>();
        d.yint      = C.yint.as<double>();
        d.lcoef     = C.lcoef.as<double>();
        d.cbase     = C.cbase.as<double>();
        d.cdelta    = C.cdelta.as<double>();
        d.nodep     = C.nodep.as<NodeP>();
        d.tnode     = C.tnode.as<double>();
        d.dtnode    = C.dtnode.as<double>();
        d.steps     = C.steps.as<uint32_t>();
        d.attempts  = C.attempts.as<uint32_t>();
        d.estructs  = C.estructs.as<EdgeStructDev>();
        if (!C.order.empty()) {
            C.node_order.upload(C.order, stream);
            device_bytes += C.node_order.bytes;
        }
        d.node_order = C.order.empty() ? nullptr : C.node_order.as<int32_t>();
        d.fx_rec    = C.fx_rec.as<FxRec>();
        d.fx_pool   = C.fx_pool.as<double>();
        d.fx_dirty  = C.fx_dirty.as<int32_t>();
        d.npi          = C.npi.nodes0.empty() ? nullptr : C.npi.nodes.as<NpiNode>();
        d.npi_thr      = C.npi.thr.as<double>();
        d.npi_thr_damp = C.npi.thr_damp.as<int32_t>();
        d.npi_damp     = C.npi.damp.as<double>();
        d.out_npi      = C.oute.nodes0.empty() ? nullptr : C.oute.nodes.as<NpiNode>();
        d.out_thr      = C.oute.thr.as<double>();
        d.out_thr_damp = C.oute.thr_damp.as<int32_t>();
        d.out_damp     = C.oute.damp.as<double>();
        d.out_d        = C.out_d.as<double>();
        C.seg_host = d;
        std::vector<SegDev> segs{d};
        std::vector<int32_t> bseg, bfirst;
        block_tables({C.num_nodes}, G, bseg, bfirst);
        up(C.seg_dev, segs);
        up(C.block_seg, bseg);
        up(C.block_first, bfirst);
        LaunchDev& ld  = C.launch;
        ld.segs        = C.seg_dev.as<SegDev>();
        ld.block_seg   = C.block_seg.as<int32_t>();
        ld.block_first = C.block_first.as<int32_t>();
        ld.inst_ids    = C.inst_ids.as<int32_t>();
        ld.num_inst    = int(C.insts.size());
        ld.num_blocks  = int(bseg.size());
        ld.G           = G;
        ld.kind        = C.kind;
        ld.npi         = (C.npi.nodes0.empty() && C.oute.nodes0.empty()) ? 0 : 1;
        ld.cyc         = cyc.as<int32_t>();
    }

    // launch tables of a cycle: one per class. (Batches keep per-G kernels so that a run's arithmetic is
    // the same compiled code alone and in any batch; mixed tables serve the request loop.)
    void launch_tables(std::vector<const LaunchDev*>& node_ld, std::vector<const LaunchDev*>& edge_ld) const
    {
        node_ld.clear();
        edge_ld.clear();
        for (const ClassHost& C : classes) {
            node_ld.push_back(&C.launch);
            bool any_edges = false;
            for (int ii : C.insts) {
                any_edges = any_edges || inst[size_t(ii)].has_edges;
            }
            if (any_edges) {
                edge_ld.push_back(&C.launch);
            }
        }
    }

    // enqueue one half-day cycle for every instance; with ev != nullptr, events bracket the node and the
    // edge launches (ev[0..1] node, ev[2..3] edge)
    void enqueue_cycle(hipEvent_t* ev = nullptr)
    {
        std::vector<const LaunchDev*> node_ld, edge_ld;
        launch_tables(node_ld, edge_ld);
        InstDev* id = inst_dev.as<InstDev>();
        hip_check(launch_clock(id, int(inst.size()), cyc.as<int32_t>(), stream), "clock kernel");
        if (!ev && concurrent_classes && node_ld.size() > 1) {
            // classes share no data: each runs its node and edge kernels on its own stream
            make_side_streams(node_ld.size());
            hip_check(hipEventRecord(fork_ev, stream), "fork");
            for (size_t k = 0; k < node_ld.size(); ++k) {
                hip_check(hipStreamWaitEvent(side[k], fork_ev, 0), "fork wait");
                hip_check(launch_node(*node_ld[k], id, side[k]), "node kernel");
                if (std::find(edge_ld.begin(), edge_ld.end(), node_ld[k]) != edge_ld.end()) {
                    hip_check(launch_edge(*node_ld[k], id, side[k]), "edge kernel");
                }
                hip_check(hipEventRecord(side_done[k], side[k]), "join");
                hip_check(hipStreamWaitEvent(stream, side_done[k], 0), "join wait");
            }
            return;
        }
        if (ev) {
            hip_check(hipEventRecord(ev[0], stream), "event");
        }
        for (const LaunchDev* ld : node_ld) {
            hip_check(launch_node(*ld, id, stream), "node kernel");
        }
        if (ev) {
            hip_check(hipEventRecord(ev[1], stream), "event");
            hip_check(hipEventRecord(ev[2], stream), "event");
        }
        for (const LaunchDev* ld : edge_ld) {
            hip_check(launch_edge(*ld, id, stream), "edge kernel");
        }
        if (ev) {
            hip_check(hipEventRecord(ev[3], stream), "event");
        }
    }

    // number of GraphSimulation loop iterations from the host mirror of every clock to its tmax
    int64_t cycles_needed(std::vector<double>* t_after = nullptr) const
    {
        int64_t mx = 0;
        for (size_t i = 0; i < inst.size(); ++i) {
            const InstHost& I = inst[i];
            double t          = I.t;
            int64_t n         = 0;
            while (t < I.tmax) {
                double step = I.dt_graph;
                if (t + step > I.tmax) {
                    step = I.tmax - t;
                }
                t = (I.single && step != I.dt_graph) ? I.tmax : t + step;
                ++n;
            }
            mx = std::max(mx, n);
            if (t_after) {
                t_after->push_back(t);
            }
        }
        return mx;
    }

    // advance every instance to its tmax (already set on host and device)
    void run(double* elapsed_ms)
    {
        std::vector<double> t_after;
        const int64_t n = cycles_needed(&t_after);
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(hipEventRecord(e0, stream), "event record");
        std::vector<hipEvent_t> evs;
        node_ms = edge_ms = 0;
        node_launches = edge_launches = 0;
        if (n > 0 && profile) {
            evs.resize(size_t(4 * n));
            for (auto& e : evs) {
                hip_check(hipEventCreate(&e), "event");
            }
            for (int64_t k = 0; k < n; ++k) {
                enqueue_cycle(&evs[size_t(4 * k)]);
            }
        }
        else if (n > 0) {
            if (use_graph) {
                if (!graph_exec) {
                    hipGraph_t g;
                    hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
                    enqueue_cycle();
                    hip_check(hipStreamEndCapture(stream, &g), "end capture");
                    hip_check(hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0), "graph instantiate");
                    (void)hipGraphDestroy(g);
                }
                for (int64_t k = 0; k < n; ++k) {
                    hip_check(hipGraphLaunch(graph_exec, stream), "graph launch");
                }
            }
            else {
                for (int64_t k = 0; k < n; ++k) {
                    enqueue_cycle();
                }
            }
        }
        hip_check(hipEventRecord(e1, stream), "event record");
        hip_check(hipEventSynchronize(e1), "run");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (size_t k = 0; k + 3 < evs.size(); k += 4) {
            float a = 0, b = 0;
            hip_check(hipEventElapsedTime(&a, evs[k], evs[k + 1]), "elapsed");
            hip_check(hipEventElapsedTime(&b, evs[k + 2], evs[k + 3]), "elapsed");
            node_ms += a;
            edge_ms += b;
        }
        if (!evs.empty()) {
            std::vector<const LaunchDev*> node_ld, edge_ld;
            launch_tables(node_ld, edge_ld);
            node_launches = int(n) * int(node_ld.size());
            edge_launches = int(n) * int(edge_ld.size());
        }
        for (auto& e : evs) {
            (void)hipEventDestroy(e);
        }
        last_ms = ms;
        if (elapsed_ms) {
            *elapsed_ms = ms;
        }
        cycles_run += n;
        for (size_t i = 0; i < inst.size(); ++i) {
            inst[i].t = t_after[i];
        }
    }

    void reset()
    {
        for (auto& C : classes) {
            hip_check(hipMemcpyAsync(C.y.p, C.y0.p, C.y.bytes, hipMemcpyDeviceToDevice, stream), "reset y");
            hip_check(hipMemcpyAsync(C.tnode.p, C.tnode0.data(), C.tnode0.size() * sizeof(double),
                                     hipMemcpyHostToDevice, stream),
                      "reset t");
            hip_check(hipMemcpyAsync(C.dtnode.p, C.dtnode0.data(), C.dtnode0.size() * sizeof(double),
                                     hipMemcpyHostToDevice, stream),
                      "reset dt");
            hip_check(hipMemsetAsync(C.steps.p, 0, C.steps.bytes, stream), "reset steps");
            hip_check(hipMemsetAsync(C.attempts.p, 0, C.attempts.bytes, stream), "reset attempts");
            hip_check(hipMemsetAsync(C.yint.p, 0xff, C.yint.bytes, stream), "reset yint");
            C.clear_fixup_records(stream);
            // damping lists, caches and NPI state as built (dynamic NPIs change them during a run)
            C.npi.reset(stream);
            C.oute.reset(stream);
            if (!C.npi.nodes0.empty()) {
                hip_check(hipMemcpyAsync(C.nodep.p, C.nodep0.data(), C.nodep0.size() * sizeof(NodeP),
                                         hipMemcpyHostToDevice, stream),
                          "reset node records");
            }
            if (C.out_d.p) {
                hip_check(hipMemsetAsync(C.out_d.p, 0, C.out_d.bytes, stream), "reset damping factors");
            }
        }
        hip_check(hipMemcpyAsync(inst_dev.p, inst0.data(), inst0.size() * sizeof(InstDev), hipMemcpyHostToDevice,
                                 stream),
                  "reset instances");
        hip_check(hipStreamSynchronize(stream), "reset");
        for (size_t i = 0; i < inst.size(); ++i) {
            inst[i].t = inst[i].t0;
        }
    }

    void set_tmax(size_t i, double tmax)
    {
        inst[i].tmax = tmax;
        InstDev* p   = inst_dev.as<InstDev>() + i;
        hip_check(hipMemcpyAsync(reinterpret_cast<char*>(p) + offsetof(InstDev, tmax), &inst[i].tmax, sizeof(double),
                                 hipMemcpyHostToDevice, stream),
                  "set tmax");
        hip_check(hipStreamSynchronize(stream), "set tmax");
    }

    std::vector<InstDev> read_inst()
    {
        std::vector<InstDev> v(inst.size());
        hip_check(hipMemcpyAsync(v.data(), inst_dev.p, v.size() * sizeof(InstDev), hipMemcpyDeviceToHost, stream),
                  "read instances");
        hip_check(hipStreamSynchronize(stream), "read instances");
        return v;
    }

    void check_status()
    {
        auto v = read_inst();
        for (size_t i = 0; i < v.size(); ++i) {
            if (v[i].status == MGO_ERR_UNSUPPORTED) {
                fail(v[i].status, "run %zu: several pending traveler moves per edge (tmax off the dt grid) are not "
                                  "supported by the device path",
                     i);
            }
            if (v[i].status != 0) {
                fail(v[i].status, "run %zu: a node made no progress within %u step attempts of one half-day", i,
                     1u << 20);
            }
        }
    }

    // daily values of instance i that are complete: days up to the current time
    int days_available(size_t i) const
    {
        const InstHost& I = inst[i];
        if (I.ndays == 0) {
            return 0;
        }
        const int last = std::min(day_floor(I.t), I.day0 + I.ndays - 1);
        return std::max(0, last - I.day0 + 1);
    }

    // sync = false leaves the copy in flight on `stream` (the caller synchronises before using `out`)
    void read_daily(size_t i, int ndays, double* out, bool sync = true)
    {
        const InstHost& I = inst[i];
        const ClassHost& C = classes[size_t(I.cls)];
        const size_t nd    = size_t(ndays) * I.num_nodes * I.G * NC;
        if (nd == 0) {
            return;
        }
        hip_check(hipMemcpyAsync(out, C.out.as<double>() + I.out_base, nd * sizeof(double), hipMemcpyDeviceToHost,
                                 stream),
                  "read daily");
        if (sync) {
            hip_check(hipStreamSynchronize(stream), "read daily");
        }
    }

    // per-node values of instance i from a segment array with `width` values per node
    template <class T>
    void gather_nodes(size_t i, const T* src, size_t width, T* out)
    {
        const InstHost& I = inst[i];
        const size_t n    = size_t(I.num_nodes);
        if (I.node_stride == 1) {
            hip_check(hipMemcpyAsync(out, src + size_t(I.node_base) * width, n * width * sizeof(T),
                                     hipMemcpyDeviceToHost, stream),
                      "read nodes");
            hip_check(hipStreamSynchronize(stream), "read nodes");
            return;
        }
        const size_t span = (n - 1) * size_t(I.node_stride) + 1;
        std::vector<T> tmp(span * width);
        hip_check(hipMemcpyAsync(tmp.data(), src + size_t(I.node_base) * width, tmp.size() * sizeof(T),
                                 hipMemcpyDeviceToHost, stream),
                  "read nodes");
        hip_check(hipStreamSynchronize(stream), "read nodes");
        for (size_t v = 0; v < n; ++v) {
            std::copy(tmp.begin() + long(v * size_t(I.node_stride) * width),
                      tmp.begin() + long((v * size_t(I.node_stride) + 1) * width), out + v * width);
        }
    }

    void read_state(size_t i, double* out)
    {
        const InstHost& I  = inst[i];
        const ClassHost& C = classes[size_t(I.cls)];
        gather_nodes(i, C.y.as<double>(), size_t(I.G) * NC, out);
    }

    void read_stats(size_t i, int64_t* steps, int64_t* attempts)
    {
        const InstHost& I  = inst[i];
        const ClassHost& C = classes[size_t(I.cls)];
        std::vector<uint32_t> s(size_t(I.num_nodes)), a(size_t(I.num_nodes));
        gather_nodes(i, C.steps.as<uint32_t>(), 1, s.data());
        gather_nodes(i, C.attempts.as<uint32_t>(), 1, a.data());
        for (size_t k = 0; k < s.size(); ++k) {
            if (steps) {
                steps[k] = s[k];
            }
            if (attempts) {
                attempts[k] = a[k];
            }
        }
    }

    void totals(int64_t& steps, int64_t& attempts)
    {
        steps = attempts = 0;
        for (auto& C : classes) {
            std::vector<uint32_t> s(size_t(C.num_nodes)), a(size_t(C.num_nodes));
            hip_check(hipMemcpyAsync(s.data(), C.steps.p, s.size() * 4, hipMemcpyDeviceToHost, stream), "stats");
            hip_check(hipMemcpyAsync(a.data(), C.attempts.p, a.size() * 4, hipMemcpyDeviceToHost, stream), "stats");
            hip_check(hipStreamSynchronize(stream), "stats");
            for (auto x : s) {
                steps += x;
            }
            for (auto x : a) {
                attempts += x;
            }
        }
    }
};

} // namespace

// ------------------------------------------------------------------------------------------ scratch
// Device scratch of the allocation-free entry point (mgo_ensemble_percentiles_noalloc): per-run offsets of
// the values and the percentile ranks the select kernel reads. One copy per GPU, held until the process
// ends; host code reaches it only through the address hipGetSymbolAddress returns for the current device.
constexpr int kScratchRuns  = 8192;
constexpr int kScratchRanks = 64;
constexpr size_t kScratchBytes = size_t(kScratchRuns) * sizeof(int64_t) + size_t(kScratchRanks) * sizeof(int32_t);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// address of this device's scratch, looked up once per device with that device current
static hipError_t scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        return hipErrorInvalidDevice;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) {
        return hipSuccess;
    }
    void* ptr        = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess) {
        g_scratch_dev[device] = ptr;
    }
    return e;
}

struct SelectRanks {
    int32_t r[kScratchRanks];
};

// per-run offsets run * positions and the ranks into the scratch (a kernel, so the call records no memset
// or copy node when captured)
__global__ void noalloc_select_setup(int64_t* base, int runs, int64_t positions, SelectRanks ranks, int num_ranks,
                                     int32_t* dranks)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < runs) {
        base[i] = int64_t(i) * positions;
    }
    if (i < num_ranks) {
        dranks[i] = ranks.r[i];
    }
}

// ------------------------------------------------------------------------------------------ context
struct ServeState;

struct mgo_context {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t hbm_budget  = 0;
    int flags          = 0;
    std::recursive_mutex mu;
    std::string last_error;
    uint64_t last_error_seq = 0; // order of this error among the context's and its serving loop's errors
    int64_t next_id = 1;
    std::unordered_map<int64_t, std::unique_ptr<Batch>> graphs;
    std::unordered_map<int64_t, std::unique_ptr<Batch>> batches;
    std::chrono::steady_clock::time_point t_origin;
    std::unique_ptr<ServeState> serve;
    int tile = 0;                                    // 0 = autotune
    std::map<std::tuple<int, int, int, int, int>, int> tile_cache; // shape -> autotuned tile
};

// numbers every recorded error so that mgo_last_error can return the newest one
static std::atomic<uint64_t> g_error_seq{0};

namespace
{
thread_local std::string tl_error;

template <class F>
mgo_status guarded(mgo_context* ctx, F&& f)
{
    if (!ctx) {
        tl_error = "context is NULL";
        return MGO_ERR_INVALID_VALUE;
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    try {
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        f();
        return MGO_OK;
    }
    catch (const Error& e) {
        ctx->last_error     = e.msg;
        ctx->last_error_seq = ++g_error_seq;
        return e.code;
    }
    catch (const std::bad_alloc&) {
        ctx->last_error     = "host allocation failed";
        ctx->last_error_seq = ++g_error_seq;
        return MGO_ERR_OUT_OF_MEMORY;
    }
}

Batch* find_batch(std::unordered_map<int64_t, std::unique_ptr<Batch>>& m, int64_t id)
{
    auto it = m.find(id);
    if (it == m.end()) {
        fail(MGO_ERR_NOT_FOUND, "unknown id %lld", (long long)id);
    }
    return it->second.get();
}
} // namespace

#include "mgo_serve.inc"

// run-time options of a batch from the context flags (read at every run)
static void apply_flags(Batch& b, int32_t flags)
{
    b.use_graph        = !(flags & MGO_CTX_NO_GRAPH_CAPTURE);
    b.profile          = (flags & MGO_CTX_PROFILE_KERNELS) != 0;
    b.sequential_fixup = (flags & MGO_CTX_SEQUENTIAL_FIXUP) != 0;
}

extern "C" {

const char* mgo_version(void)
{
    return "mgo 0.1 (gfx950, fp64)";
}

mgo_status mgo_context_create(const mgo_device_cfg* cfg, mgo_context** out)
{
    if (!out) {
        return MGO_ERR_INVALID_VALUE;
    }
    *out = nullptr;
    auto ctx      = std::make_unique<mgo_context>();
    ctx->device   = cfg ? cfg->device : 0;
    ctx->flags    = cfg ? cfg->flags : 0;
    ctx->t_origin = std::chrono::steady_clock::now();
    int n         = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= ctx->device) {
        tl_error = "no HIP device available";
        return MGO_ERR_DEVICE;
    }
    if (hipSetDevice(ctx->device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        tl_error = "cannot create HIP stream";
        return MGO_ERR_DEVICE;
    }
    if (scratch_setup(ctx->device) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        tl_error = "cannot resolve the device scratch";
        return MGO_ERR_DEVICE;
    }
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    ctx->hbm_budget = (cfg && cfg->hbm_budget_bytes) ? cfg->hbm_budget_bytes : size_t(double(free_b) * 0.9);
    *out            = ctx.release();
    return MGO_OK;
}

mgo_status mgo_context_set_flags(mgo_context* ctx, int32_t flags)
{
    return guarded(ctx, [&] {
        if (flags & ~(MGO_CTX_NO_GRAPH_CAPTURE | MGO_CTX_PROFILE_KERNELS | MGO_CTX_SEQUENTIAL_FIXUP)) {
            fail(MGO_ERR_INVALID_VALUE, "unknown context flags 0x%x", unsigned(flags));
        }
        ctx->flags = flags;
    });
}

mgo_status mgo_context_set_tile(mgo_context* ctx, int32_t tile)
{
    return guarded(ctx, [&] {
        if (tile < 0) {
            fail(MGO_ERR_INVALID_VALUE, "tile must be >= 0");
        }
        ctx->tile = tile;
    });
}

mgo_status mgo_context_destroy(mgo_context* ctx)
{
    if (!ctx) {
        return MGO_ERR_INVALID_VALUE;
    }
    serve_shutdown(ctx);
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        ctx->graphs.clear();
        ctx->batches.clear();
        if (ctx->stream) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipStreamDestroy(ctx->stream);
        }
    }
    delete ctx;
    return MGO_OK;
}

const char* mgo_last_error(const mgo_context* ctx)
{
    if (!ctx) {
        return tl_error.c_str();
    }
    tl_error = ctx->last_error;
    if (ctx->serve) {
        // the serving loop records its errors (worker, preparation, submissions) on its own; the newer wins
        std::lock_guard<std::mutex> lk(ctx->serve->mu);
        if (ctx->serve->error_seq > ctx->last_error_seq) {
            tl_error = ctx->serve->error;
        }
    }
    return tl_error.c_str();
}

double mgo_clock_ms(const mgo_context* ctx)
{
    if (!ctx) {
        return 0.0;
    }
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ctx->t_origin).count();
}

mgo_status mgo_graph_create(mgo_context* ctx, const mgo_graph_desc* graph, const mgo_sim_cfg* cfg, mgo_graph_id* out)
{
    return guarded(ctx, [&] {
        if (!cfg || !out) {
            fail(MGO_ERR_INVALID_VALUE, "cfg and out are required");
        }
        if (!(cfg->horizon >= cfg->t0)) {
            fail(MGO_ERR_INVALID_VALUE, "cfg.horizon must be >= t0 (it sizes the daily output)");
        }
        std::vector<RunSpec> runs(1);
        runs[0] = RunSpec{graph, *cfg, cfg->horizon, false, {}};
        auto b  = std::make_unique<Batch>();
        apply_flags(*b, ctx->flags);
        b->build(runs, ctx->stream);
        *out = ctx->next_id++;
        ctx->graphs[*out] = std::move(b);
    });
}

mgo_status mgo_graph_advance(mgo_context* ctx, mgo_graph_id id, double tmax)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->graphs, id);
        if (!std::isfinite(tmax)) {
            fail(MGO_ERR_INVALID_VALUE, "tmax must be finite");
        }
        if (day_floor(tmax) > b->inst[0].day0 + b->inst[0].ndays - 1 && b->inst[0].ndays >= 0) {
            fail(MGO_ERR_OUT_OF_RANGE, "tmax %g beyond the horizon given at creation", tmax);
        }
        apply_flags(*b, ctx->flags);
        b->set_tmax(0, tmax);
        b->run(nullptr);
        b->check_status();
    });
}

mgo_status mgo_graph_get_time(mgo_context* ctx, mgo_graph_id id, double* t)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->graphs, id);
        if (!t) {
            fail(MGO_ERR_INVALID_VALUE, "t is NULL");
        }
        *t = b->inst[0].t;
    });
}

mgo_status mgo_graph_read_state(mgo_context* ctx, mgo_graph_id id, double* y_out)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->graphs, id);
        if (!y_out) {
            fail(MGO_ERR_INVALID_VALUE, "y_out is NULL");
        }
        b->read_state(0, y_out);
    });
}

mgo_status mgo_graph_read_daily(mgo_context* ctx, mgo_graph_id id, int32_t capacity_days, double* out, int32_t* num_days)
{
    return guarded(ctx, [&] {
        Batch* b  = find_batch(ctx->graphs, id);
        const int n = std::min<int>(b->days_available(0), capacity_days);
        if (out && n > 0) {
            b->read_daily(0, n, out);
        }
        if (num_days) {
            *num_days = b->days_available(0);
        }
    });
}

mgo_status mgo_graph_read_stats(mgo_context* ctx, mgo_graph_id id, int64_t* steps, int64_t* attempts)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->graphs, id);
        b->read_stats(0, steps, attempts);
    });
}

mgo_status mgo_graph_destroy(mgo_context* ctx, mgo_graph_id id)
{
    return guarded(ctx, [&] {
        find_batch(ctx->graphs, id);
        ctx->graphs.erase(id);
    });
}

mgo_status mgo_simulate_single(mgo_context* ctx, const mgo_graph_desc* models, const mgo_sim_cfg* cfg, double tmax,
                               int32_t capacity_days, double* daily_out, int32_t* num_days)
{
    return guarded(ctx, [&] {
        if (!cfg || !models) {
            fail(MGO_ERR_INVALID_VALUE, "models and cfg are required");
        }
        if (models->num_edges != 0) {
            fail(MGO_ERR_INVALID_VALUE, "mgo_simulate_single takes independent models (num_edges == 0)");
        }
        if (!(tmax > cfg->t0)) {
            fail(MGO_ERR_INVALID_VALUE, "tmax must be > t0");
        }
        std::vector<RunSpec> runs(1);
        runs[0] = RunSpec{models, *cfg, tmax, true, {}};
        Batch b;
        b.use_graph = false;
        b.build(runs, ctx->stream);
        b.set_tmax(0, tmax);
        b.run(nullptr);
        b.check_status();
        const int n = std::min<int>(b.days_available(0), capacity_days);
        if (daily_out && n > 0) {
            b.read_daily(0, n, daily_out);
        }
        if (num_days) {
            *num_days = b.days_available(0);
        }
    });
}

// tmax of every run on the host mirror and in the device table; the table becomes the reset state
static void set_run_tmax(Batch& b, const std::vector<double>& tmax, hipStream_t stream)
{
    std::vector<InstDev> v = b.read_inst();
    for (size_t r = 0; r < tmax.size(); ++r) {
        b.inst[r].tmax = tmax[r];
        v[r].tmax      = tmax[r];
    }
    hip_check(hipMemcpyAsync(b.inst_dev.p, v.data(), v.size() * sizeof(InstDev), hipMemcpyHostToDevice, stream),
              "tmax upload");
    hip_check(hipStreamSynchronize(stream), "tmax upload");
    b.inst0 = v;
}

// Tile autotuner. Takes the largest group of runs that share a graph; if it has at least four runs, each
// candidate tile is timed on a trial batch of up to 256 of those runs over two simulated days (direct
// launches, one warm-up run, then the faster of two timed runs) and the fastest is kept for this shape (number of age
// groups, integrator, nodes, edges, group size bucket) for the life of the context.
static int autotune_tile(mgo_context* ctx, const std::vector<RunSpec>& runs)
{
    if (ctx->tile > 0) {
        return ctx->tile;
    }
    std::map<std::tuple<const void*, const void*, const void*, int, int, int, int>, std::vector<size_t>> groups;
    for (size_t r = 0; r < runs.size(); ++r) {
        const auto* g = runs[r].g;
        groups[std::make_tuple((const void*)g->edge_start, (const void*)g->edge_end, (const void*)g->edge_coeffs,
                               g->num_edges, g->num_nodes, g->num_groups, runs[r].cfg.integrator.kind)]
            .push_back(r);
    }
    const std::vector<size_t>* best = nullptr;
    for (auto& kv : groups) {
        if (!best || kv.second.size() > best->size()) {
            best = &kv.second;
        }
    }
    if (!best) {
        return 1;
    }
    const RunSpec& r0 = runs[(*best)[0]];
    const int bucket  = best->size() >= 256 ? 256 : int(best->size());
    if (bucket < 4) {
        return 1;
    }
    const auto key    = std::make_tuple(r0.g->num_groups, r0.cfg.integrator.kind, r0.g->num_nodes, r0.g->num_edges, bucket);
    auto hit          = ctx->tile_cache.find(key);
    if (hit != ctx->tile_cache.end()) {
        return hit->second;
    }
    std::vector<RunSpec> trial;
    std::vector<double> tmax;
    for (int k = 0; k < bucket; ++k) {
        RunSpec t = runs[(*best)[size_t(k)]];
        t.tmax    = t.cfg.t0 + 4.0 * (t.cfg.dt_graph > 0 ? t.cfg.dt_graph : 0.5);
        trial.push_back(t);
        tmax.push_back(t.tmax);
    }
    int best_tile  = 1;
    double best_ms = INFINITY;
    // powers of two, and the wave-aligned tiles: 64 / G runs put exactly one graph node in each wavefront
    // of the edge kernels (all lanes walk the same edge list), twice that pairs two such waves per tile
    const int npw = WAVE / std::max(1, r0.g->num_groups);
    std::vector<int> cands = {1, 2, 4, 8, 16, 64, npw, 2 * npw};
    std::sort(cands.begin(), cands.end());
    cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
    for (int c : cands) {
        if (c > bucket) {
            break;
        }
        Batch tb;
        tb.use_graph = false;
        tb.build(trial, ctx->stream, c);
        set_run_tmax(tb, tmax, ctx->stream);
        double ms = 0, ms2 = 0;
        tb.run(&ms); // warm-up
        tb.reset();
        tb.run(&ms);
        tb.reset();
        tb.run(&ms2);
        ms = std::min(ms, ms2); // the faster of two timed runs: one slow run does not decide
        if (ms < best_ms) {
            best_ms   = ms;
            best_tile = c;
        }
    }
    ctx->tile_cache[key] = best_tile;
    return best_tile;
}

static void build_ensemble(mgo_context* ctx, const mgo_ensemble_desc* desc, Batch& b)
{
    if (!desc || desc->num_runs < 1 || !desc->graphs || !desc->cfgs || !desc->tmax) {
        fail(MGO_ERR_INVALID_VALUE, "ensemble description incomplete");
    }
    std::vector<RunSpec> runs(size_t(desc->num_runs));
    for (int r = 0; r < desc->num_runs; ++r) {
        if (!std::isfinite(desc->tmax[r]) || !(desc->tmax[r] >= desc->cfgs[r].t0)) {
            fail(MGO_ERR_INVALID_VALUE, "run %d: tmax must be finite and >= t0", r);
        }
        runs[size_t(r)] = RunSpec{&desc->graphs[r], desc->cfgs[r], desc->tmax[r], false, {}};
    }
    const int tile = autotune_tile(ctx, runs);
    apply_flags(b, ctx->flags);
    b.build(runs, ctx->stream, tile);
    set_run_tmax(b, std::vector<double>(desc->tmax, desc->tmax + desc->num_runs), ctx->stream);
}

static void fill_info(Batch& b, mgo_run_info* info, double ms, int64_t cycles)
{
    if (!info) {
        return;
    }
    std::memset(info, 0, sizeof(*info));
    info->device_ms = ms;
    double nd       = 0;
    for (auto& I : b.inst) {
        nd += double(I.num_nodes) * (I.tmax - I.t0);
    }
    info->node_days = nd;
    info->tile      = b.tile;
    b.totals(info->steps, info->attempts);
    auto v = b.read_inst();
    for (auto& I : v) {
        info->exact_replays += I.replays;
        info->fixup_passes += I.fx_passes;
    }
    info->half_day_cycles = int32_t(cycles);
    info->node_launches   = b.node_launches;
    info->node_kernel_ms  = b.node_ms;
    info->edge_launches   = b.edge_launches;
    info->edge_kernel_ms  = b.edge_ms;
}

static void read_all_daily(Batch& b, double* out, size_t first = 0, size_t count = SIZE_MAX)
{
    size_t off = 0;
    for (size_t i = first; i < b.inst.size() && i - first < count; ++i) {
        const auto& I = b.inst[i];
        b.read_daily(i, I.ndays, out + off);
        off += size_t(I.ndays) * I.num_nodes * I.G * NC;
    }
}

mgo_status mgo_ensemble_run(mgo_context* ctx, const mgo_ensemble_desc* desc, double* daily_out, mgo_run_info* info)
{
    return guarded(ctx, [&] {
        Batch b;
        build_ensemble(ctx, desc, b);
        const int64_t cyc = b.cycles_needed();
        double ms         = 0;
        b.run(&ms);
        b.check_status();
        if (daily_out) {
            read_all_daily(b, daily_out);
        }
        fill_info(b, info, ms, cyc);
    });
}

mgo_status mgo_batch_create(mgo_context* ctx, const mgo_ensemble_desc* desc, mgo_batch_id* out)
{
    return guarded(ctx, [&] {
        if (!out) {
            fail(MGO_ERR_INVALID_VALUE, "out is NULL");
        }
        auto b = std::make_unique<Batch>();
        build_ensemble(ctx, desc, *b);
        *out = ctx->next_id++;
        ctx->batches[*out] = std::move(b);
    });
}

mgo_status mgo_batch_run(mgo_context* ctx, mgo_batch_id id, mgo_run_info* info)
{
    return guarded(ctx, [&] {
        Batch* b          = find_batch(ctx->batches, id);
        apply_flags(*b, ctx->flags);
        const int64_t cyc = b->cycles_needed();
        double ms         = 0;
        b->run(&ms);
        b->check_status();
        fill_info(*b, info, ms, cyc);
    });
}

mgo_status mgo_batch_read_daily(mgo_context* ctx, mgo_batch_id id, int32_t first_run, int32_t num_runs,
                                double* daily_out)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->batches, id);
        if (!daily_out) {
            fail(MGO_ERR_INVALID_VALUE, "daily_out is NULL");
        }
        if (first_run < 0 || num_runs < 0 || size_t(first_run) + size_t(num_runs) > b->inst.size()) {
            fail(MGO_ERR_OUT_OF_RANGE, "runs %d..%d out of range", first_run, first_run + num_runs);
        }
        read_all_daily(*b, daily_out, size_t(first_run), size_t(num_runs));
    });
}

// Runs [first, first + num) of a batch as one ensemble: same class, node count, age groups and days.
// Fills the per-run output offsets (doubles into the class's output buffer) on the device.
struct EnsembleView {
    const double* out = nullptr;
    int nodes = 0, G = 0, days = 0;
    DevBuf base;
};

static void ensemble_view(Batch& b, int32_t first, int32_t num, EnsembleView& v)
{
    if (first < 0 || num < 1 || size_t(first) + size_t(num) > b.inst.size()) {
        fail(MGO_ERR_OUT_OF_RANGE, "runs %d..%d outside the batch", first, first + num - 1);
    }
    const InstHost& I0 = b.inst[size_t(first)];
    std::vector<int64_t> base;
    for (int r = first; r < first + num; ++r) {
        const InstHost& I = b.inst[size_t(r)];
        if (I.cls != I0.cls || I.num_nodes != I0.num_nodes || I.ndays != I0.ndays || I.day0 != I0.day0) {
            fail(MGO_ERR_INVALID_VALUE, "run %d differs in shape from run %d (nodes, age groups, days)", r, first);
        }
        base.push_back(I.out_base);
    }
    if (I0.ndays < 1) {
        fail(MGO_ERR_INVALID_VALUE, "runs have no daily output");
    }
    v.out   = b.classes[size_t(I0.cls)].out.as<double>();
    v.nodes = I0.num_nodes;
    v.G     = I0.G;
    v.days  = I0.ndays;
    v.base.upload(base, b.stream);
}

static void download(Batch& b, const DevBuf& src, size_t doubles, double* dst)
{
    hip_check(hipMemcpyAsync(dst, src.p, doubles * sizeof(double), hipMemcpyDeviceToHost, b.stream), "download");
    hip_check(hipStreamSynchronize(b.stream), "download");
}

// percentile ranks floor(num_runs * p) (analyze_result.h:339), checked
static std::vector<int32_t> percentile_ranks(int32_t num_runs, const double* p, int32_t num_p)
{
    if (num_runs > 8192) {
        fail(MGO_ERR_UNSUPPORTED, "percentiles over more than 8192 runs");
    }
    std::vector<int32_t> ranks;
    for (int q = 0; q < num_p; ++q) {
        if (!(p[q] > 0.0 && p[q] < 1.0)) {
            fail(MGO_ERR_INVALID_VALUE, "percentile %g outside (0, 1)", p[q]);
        }
        ranks.push_back(int32_t(static_cast<size_t>(double(num_runs) * p[q])));
    }
    return ranks;
}

mgo_status mgo_batch_ensemble_mean(mgo_context* ctx, mgo_batch_id id, int32_t first_run, int32_t num_runs,
                                   double* mean_out)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->batches, id);
        if (!mean_out) {
            fail(MGO_ERR_INVALID_VALUE, "output is NULL");
        }
        EnsembleView v;
        ensemble_view(*b, first_run, num_runs, v);
        const int64_t positions = int64_t(v.days) * v.nodes * v.G * NC;
        DevBuf mean;
        mean.alloc(size_t(positions) * sizeof(double));
        hip_check(launch_ensemble_mean(v.out, v.base.as<int64_t>(), num_runs, positions, mean.as<double>(), b->stream),
                  "ensemble mean");
        download(*b, mean, size_t(positions), mean_out);
    });
}

mgo_status mgo_batch_ensemble_percentiles(mgo_context* ctx, mgo_batch_id id, int32_t first_run, int32_t num_runs,
                                          const double* p, int32_t num_p, double* out)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->batches, id);
        if (!out || !p || num_p < 1) {
            fail(MGO_ERR_INVALID_VALUE, "percentiles or output missing");
        }
        const std::vector<int32_t> ranks = percentile_ranks(num_runs, p, num_p);
        EnsembleView v;
        ensemble_view(*b, first_run, num_runs, v);
        const int64_t positions = int64_t(v.days) * v.nodes * v.G * NC;
        DevBuf dranks, sel;
        dranks.upload(ranks, b->stream);
        sel.alloc(size_t(positions) * size_t(num_p) * sizeof(double));
        hip_check(launch_ensemble_select(v.out, v.base.as<int64_t>(), num_runs, positions, dranks.as<int32_t>(), num_p,
                                         sel.as<double>(), b->stream),
                  "ensemble percentiles");
        download(*b, sel, size_t(positions) * size_t(num_p), out);
    });
}

// host values [run][position] on the device, with the per-run offsets the ensemble kernels take
static void upload_values(mgo_context* ctx, const double* values, int32_t num_runs, int64_t positions, DevBuf& dv,
                          DevBuf& base)
{
    if (!values || num_runs < 1 || positions < 1) {
        fail(MGO_ERR_INVALID_VALUE, "values missing or empty (%d runs, %lld positions)", num_runs,
             (long long)positions);
    }
    const size_t n = size_t(num_runs) * size_t(positions);
    dv.alloc(n * sizeof(double));
    hip_check(hipMemcpyAsync(dv.p, values, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream), "upload values");
    std::vector<int64_t> b(static_cast<size_t>(num_runs));
    for (int32_t r = 0; r < num_runs; ++r) {
        b[size_t(r)] = int64_t(r) * positions;
    }
    base.upload(b, ctx->stream);
}

static void download_values(mgo_context* ctx, const DevBuf& src, size_t doubles, double* dst)
{
    hip_check(hipMemcpyAsync(dst, src.p, doubles * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "download");
    hip_check(hipStreamSynchronize(ctx->stream), "download");
}

mgo_status mgo_ensemble_mean_values(mgo_context* ctx, const double* values, int32_t num_runs, int64_t num_positions,
                                    double* mean_out)
{
    return guarded(ctx, [&] {
        if (!mean_out) {
            fail(MGO_ERR_INVALID_VALUE, "output is NULL");
        }
        DevBuf dv, base, mean;
        upload_values(ctx, values, num_runs, num_positions, dv, base);
        mean.alloc(size_t(num_positions) * sizeof(double));
        hip_check(launch_ensemble_mean(dv.as<double>(), base.as<int64_t>(), num_runs, num_positions,
                                       mean.as<double>(), ctx->stream),
                  "ensemble mean");
        download_values(ctx, mean, size_t(num_positions), mean_out);
    });
}

mgo_status mgo_ensemble_percentiles_values(mgo_context* ctx, const double* values, int32_t num_runs,
                                           int64_t num_positions, const double* p, int32_t num_p, double* out)
{
    return guarded(ctx, [&] {
        if (!out || !p || num_p < 1) {
            fail(MGO_ERR_INVALID_VALUE, "percentiles or output missing");
        }
        const std::vector<int32_t> ranks = percentile_ranks(num_runs, p, num_p);
        DevBuf dv, base, dranks, sel;
        upload_values(ctx, values, num_runs, num_positions, dv, base);
        dranks.upload(ranks, ctx->stream);
        sel.alloc(size_t(num_positions) * size_t(num_p) * sizeof(double));
        hip_check(launch_ensemble_select(dv.as<double>(), base.as<int64_t>(), num_runs, num_positions,
                                         dranks.as<int32_t>(), num_p, sel.as<double>(), ctx->stream),
                  "ensemble percentiles");
        download_values(ctx, sel, size_t(num_positions) * size_t(num_p), out);
    });
}

mgo_status mgo_ensemble_percentiles_noalloc(mgo_context* ctx, const double* values_dev, int32_t num_runs,
                                            int64_t num_positions, const double* p, int32_t num_p, double* out_dev,
                                            void* stream)
{
    return guarded(ctx, [&] {
        if (!values_dev || !out_dev || !p || num_p < 1 || num_runs < 1 || num_positions < 1) {
            fail(MGO_ERR_INVALID_VALUE, "values, percentiles or output missing or empty");
        }
        if (num_runs > kScratchRuns || num_p > kScratchRanks) {
            fail(MGO_ERR_OUT_OF_RANGE, "%d runs / %d percentiles exceed the scratch of %d runs / %d percentiles",
                 num_runs, num_p, kScratchRuns, kScratchRanks);
        }
        SelectRanks rk{};
        for (int q = 0; q < num_p; ++q) {
            if (!(p[q] > 0.0 && p[q] < 1.0)) {
                fail(MGO_ERR_INVALID_VALUE, "percentile %g outside (0, 1)", p[q]);
            }
            rk.r[q] = int32_t(static_cast<size_t>(double(num_runs) * p[q])); // analyze_result.h:339
        }
        void* scr = g_scratch_dev[ctx->device];
        if (!scr) {
            fail(MGO_ERR_DEVICE, "device scratch not resolved");
        }
        auto* base     = static_cast<int64_t*>(scr);
        auto* dranks   = reinterpret_cast<int32_t*>(static_cast<unsigned char*>(scr) + kScratchRuns * sizeof(int64_t));
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        const int n    = std::max(num_runs, num_p);
        hipLaunchKernelGGL(noalloc_select_setup, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, base, num_runs,
                           num_positions, rk, num_p, dranks);
        hip_check(hipGetLastError(), "scratch setup");
        hip_check(launch_ensemble_select(values_dev, base, num_runs, num_positions, dranks, num_p, out_dev, st),
                  "ensemble percentiles");
        if (!stream) {
            hip_check(hipStreamSynchronize(st), "ensemble percentiles");
        }
    });
}

mgo_status mgo_sum_nodes_values(mgo_context* ctx, const double* values, int32_t num_runs, int32_t num_days,
                                int32_t num_nodes, int32_t num_elements, double* out)
{
    return guarded(ctx, [&] {
        if (!out) {
            fail(MGO_ERR_INVALID_VALUE, "output is NULL");
        }
        if (num_days < 1 || num_nodes < 1 || num_elements < 1) {
            fail(MGO_ERR_INVALID_VALUE, "days, nodes and elements must be positive");
        }
        const int64_t per_run = int64_t(num_days) * num_nodes * num_elements;
        DevBuf dv, base, sums;
        upload_values(ctx, values, num_runs, per_run, dv, base);
        const size_t n = size_t(num_runs) * size_t(num_days) * size_t(num_elements);
        sums.alloc(n * sizeof(double));
        hip_check(launch_sum_nodes(dv.as<double>(), base.as<int64_t>(), num_runs, num_days, num_nodes, num_elements,
                                   sums.as<double>(), ctx->stream),
                  "sum nodes");
        download_values(ctx, sums, n, out);
    });
}

mgo_status mgo_batch_sum_nodes(mgo_context* ctx, mgo_batch_id id, int32_t first_run, int32_t num_runs, double* out)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->batches, id);
        if (!out) {
            fail(MGO_ERR_INVALID_VALUE, "output is NULL");
        }
        EnsembleView v;
        ensemble_view(*b, first_run, num_runs, v);
        const int width = v.G * NC;
        const size_t n  = size_t(num_runs) * size_t(v.days) * size_t(width);
        DevBuf sums;
        sums.alloc(n * sizeof(double));
        hip_check(launch_sum_nodes(v.out, v.base.as<int64_t>(), num_runs, v.days, v.nodes, width, sums.as<double>(),
                                   b->stream),
                  "sum nodes");
        download(*b, sums, n, out);
    });
}

mgo_status mgo_batch_reset(mgo_context* ctx, mgo_batch_id id)
{
    return guarded(ctx, [&] {
        Batch* b = find_batch(ctx->batches, id);
        b->reset();
    });
}

mgo_status mgo_batch_destroy(mgo_context* ctx, mgo_batch_id id)
{
    return guarded(ctx, [&] {
        find_batch(ctx->batches, id);
        ctx->batches.erase(id);
    });
}

} // extern "C"
