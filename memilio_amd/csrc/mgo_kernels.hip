// mgo_kernels.hip - gfx950 kernels of the Graph-ODE engine (FP64 throughout).
//
// One lane per (node, age group): a wavefront holds 64/G nodes; a node's 10 compartments of one age
// group live in the lane's registers. Cross-age-group terms (force of infection, test-and-trace demand,
// ICU occupancy, the step's error norm) are group reductions over the node's G consecutive lanes with
// ds_bpermute shuffles; the G x G contact row of the lane stays in registers. There is no dense
// contraction, so everything runs on the FP64 vector ALUs.
//
// Kernels per half-day cycle (GraphSimulation::advance, cpp/memilio/mobility/graph_simulation.h:101-120):
//   clock_kernel       per instance: step size, edge action and output day of this cycle
//   node_phase<G,K>    per node lane: SystemIntegrator::advance to t+dt with integrator K; writes the
//                      departure row (departure cycles) or the pre-return state and the return
//                      coefficients (return cycles)
//   depart_phase<G>    per node lane: the node's departures over all incident edges, walked in the
//                      reference's edge order (deterministic segmented sum, no atomics)
//   return_phase<G>    per node lane: the returns with the negative-value fix-up, in parallel passes over
//                      the nodes whose inputs changed (fixup_check between passes)
//   replay_phase<G>    per flagged instance: sequential return pass with the negative-value fix-up
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "../../include/mgo.h"
#include "mgo_internal.h"
#include "mgo_npi.h"

namespace mgo
{

#define MGO_PI 3.14159265358979323846

// cpp/memilio/math/smoother.h:46-60
__device__ __forceinline__ double smoother_cosine(double x, double xl, double xr, double yl, double yr)
{
    if (x <= xl) {
        return yl;
    }
    if (x >= xr) {
        return yr;
    }
    return 0.5 * (yl - yr) * cos(MGO_PI / (xr - xl) * (x - xl)) + 0.5 * (yl + yr);
}

// Blocks are dispatched round-robin over the 8 XCDs; map the hardware block id so that each XCD gets a
// contiguous range of nodes (instances stay within one XCD's L2 for the edge gathers). The host pads
// the grid to a multiple of 8.
__device__ __forceinline__ int xcd_block()
{
    const int per = int(gridDim.x) >> 3;
    const int b   = int(blockIdx.x);
    return (b & 7) * per + (b >> 3);
}

template <int G>
__device__ __forceinline__ double group_sum(double v, int base)
{
    double s = __shfl(v, base);
#pragma unroll
    for (int j = 1; j < G; ++j) {
        s = s + __shfl(v, base + j);
    }
    return s;
}

template <int G>
__device__ __forceinline__ double group_max(double v, int base)
{
    double s = __shfl(v, base);
#pragma unroll
    for (int j = 1; j < G; ++j) {
        s = fmax(s, __shfl(v, base + j));
    }
    return s;
}

template <int G>
__device__ __forceinline__ double group_min(double v, int base)
{
    double s = __shfl(v, base);
#pragma unroll
    for (int j = 1; j < G; ++j) {
        s = fmin(s, __shfl(v, base + j));
    }
    return s;
}

// predicate over the node's G lanes from one wavefront ballot (lanes of a node are active together)
template <int G>
__device__ __forceinline__ bool group_all(bool v, int base)
{
    const unsigned long long m = (~__ballot(!v) >> base) & ((1ull << G) - 1);
    return m == ((1ull << G) - 1);
}

// x^p for the step-size controllers' factors (x > 0, |p| < 1): exp2(p log2 x) within a few ulp of pow, at
// about half of its cost (pow carries log(x) in extended precision for arbitrary exponents)
__device__ __forceinline__ double step_pow(double x, double p)
{
    return exp2(p * log2(x));
}

// 1/x for a population x >= 1e-12: the hardware reciprocal refined by two Newton steps (within an ulp of the
// correctly rounded quotient, half the instructions of the IEEE division sequence)
__device__ __forceinline__ double recip_pop(double x)
{
    double r       = __builtin_amdgcn_rcp(x);
    double e       = fma(-x, r, 1.0);
    r              = fma(r, e, r);
    e              = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Per-lane context of one node lane, loaded once per kernel.
template <int G>
struct Lane {
    double c[NUM_LC];
    double crow[G];   // contact row at the last evaluated time (constant without dampings)
    NodeP np;
    int node;
    int age;
    int base;         // first lane of the node's group
    double* wx;       // LDS: the wavefront's exchange slots, one double per lane (force of infection, error norm)
    double* ws;       // LDS: the wavefront's stage-season slots, (5 + G - 1) / G planes of BLOCK doubles
};

// Effective contact row C(t)_age,. = B - D(t) o (B - M) (contact_matrix.h:190-194, damping.h:381-396)
template <int G>
__device__ __forceinline__ void contact_row(const SegDev& c, Lane<G>& L, double t)
{
    const double* B = c.cbase + (size_t(L.node) * G + L.age) * G;
    if (!(L.np.flags & NF_DAMPED)) {
#pragma unroll
        for (int j = 0; j < G; ++j) {
            L.crow[j] = B[j];
        }
        return;
    }
    const double* tk = L.np.damp_time;
    int k            = 1;
    while (k < L.np.damp_cnt - 1 && !(tk[k] > t)) {
        ++k;
    }
    const double xr    = tk[k];
    const double xl    = xr - 1.0;
    const double* Dl   = L.np.damp_mat + (size_t(k - 1) * G + L.age) * G;
    const double* Dr   = L.np.damp_mat + (size_t(k) * G + L.age) * G;
    const double* Dd   = c.cdelta + (size_t(L.node) * G + L.age) * G;
    const bool left    = t <= xl;
    const bool right   = t >= xr;
    const double cs    = (left || right) ? 0.0 : cos(MGO_PI / (xr - xl) * (t - xl));
#pragma unroll
    for (int j = 0; j < G; ++j) {
        const double d = left ? Dl[j] : (right ? Dr[j] : 0.5 * (Dl[j] - Dr[j]) * cs + 0.5 * (Dl[j] + Dr[j]));
        L.crow[j]      = B[j] - d * Dd[j];
    }
}

// Seasonal factor 1 + k sin(pi ((StartDay + t) / 182.5 + 0.5)) of the lane's node at time t
// (model.h:115-117).
template <int G>
__device__ __forceinline__ double season_at(const Lane<G>& L, double t)
{
    return 1.0 + L.np.seasonality * sin(MGO_PI * ((L.np.start_day + t) / 182.5 + 0.5));
}

// Seasonal factors of the five later stages of a step, spread over the node's G lanes: lane a evaluates
// stages j = a, a + G, ... (j < 5) at t + c_j h, and every lane reads a stage's factor from the lane that
// owns it (one sin per lane and step attempt for G >= 5 instead of five).
template <int G>
struct StageSeasons {
    static constexpr int N = (5 + G - 1) / G;
    double v[N];
};

template <int G>
__device__ __forceinline__ StageSeasons<G> stage_seasons(const Lane<G>& L, double t, double h, double c2, double c3,
                                                         double c4, double c5, double c6)
{
    StageSeasons<G> s;
#pragma unroll
    for (int q = 0; q < StageSeasons<G>::N; ++q) {
        const int j    = L.age + q * G;
        const double c = (j == 0) ? c2 : (j == 1) ? c3 : (j == 2) ? c4 : (j == 3) ? c5 : c6;
        s.v[q]         = (j < 5) ? season_at<G>(L, t + c * h) : 1.0;
        L.ws[q * BLOCK + threadIdx.x % WAVE] = s.v[q]; // read by the node's lanes from the wavefront's slots
    }
    __builtin_amdgcn_wave_barrier();
    return s;
}

template <int G, int J>
__device__ __forceinline__ double stage_season(const Lane<G>& L, const StageSeasons<G>&)
{
    return L.ws[(J / G) * BLOCK + L.base + J % G];
}

// Force-of-infection coefficient of the lane's age group: season * tp * sum_j C_gj(t) w_j(pop),
// w_j = divN_j * (rel_j pop[INS_j] + risk_j pop[ISy_j]) (model.h:99-145, factored over j). Also returns
// the ICU-adjusted ISev->ICr / ISev->D coefficients (model.h:147-152, 196-204).
template <int G>
__device__ __forceinline__ double force_of_infection(const SegDev& c, Lane<G>& L, const double (&pop)[NC], double t,
                                                     double season, double& c_icr, double& c_d)
{
    double risk = L.c[LC_RISK];
    if (L.np.flags & NF_TTC_ACTIVE) {
        const double ttr = group_sum<G>(L.c[LC_TTR] * pop[INS], L.base);
        risk             = smoother_cosine(ttr, L.np.ttc, L.np.ttc_hi, L.c[LC_RISK], L.c[LC_MAXRISK]);
    }
    const double Nj   = pop[S] + pop[E] + pop[INS] + pop[ISY] + pop[ISEV] + pop[ICR] + pop[R];
    const double divN = (Nj < 1e-12) ? 0.0 : recip_pop(Nj);
    const double w    = divN * (L.c[LC_REL] * pop[INS] + risk * pop[ISY]);
    if (L.np.flags & NF_DAMPED) {
        contact_row<G>(c, L, t);
    }
    // w of the node's G lanes through the wavefront's LDS slots: one write, G reads (a wavefront's LDS
    // instructions complete in order; the lanes of a node are active together)
    L.wx[threadIdx.x % WAVE] = w;
    __builtin_amdgcn_wave_barrier();
    double lam = 0.0;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        lam += L.crow[j] * L.wx[L.base + j];
    }
    __builtin_amdgcn_wave_barrier();
    c_icr = L.c[LC_ISEV_ICR];
    c_d   = L.c[LC_ISEV_D];
    if (L.np.flags & NF_ICU_ACTIVE) {
        const double icu  = group_sum<G>(pop[ICR], L.base);
        const double crit = smoother_cosine(icu, 0.90 * L.np.icu, L.np.icu, L.c[LC_CPS], 0.0);
        c_icr             = crit / L.c[LC_TISEV];
        c_d               = (L.c[LC_DPS] + (L.c[LC_CPS] - crit)) / L.c[LC_TISEV];
    }
    return season * L.c[LC_TP] * lam;
}

// flows of one age group -> derivative (flow_model.h:112-125,219-235 with the osecir Flows order)
__device__ __forceinline__ void flows_to_dydt(const double (&y)[NC], double kappa, double cE, double cINS_ISY,
                                              double cINS_R, double cISY_ISEV, double cISY_R, double cISEV_ICR,
                                              double cISEV_R, double cISEV_D, double cICR_D, double cICR_R,
                                              double (&dy)[NC])
{
    const double f0  = y[S] * kappa;
    const double f1  = cE * y[E];
    const double f2  = cINS_ISY * y[INS];
    const double f3  = cINS_R * y[INS];
    const double f4  = cINS_ISY * y[INSC];
    const double f5  = cINS_R * y[INSC];
    const double f6  = cISY_ISEV * y[ISY];
    const double f7  = cISY_R * y[ISY];
    const double f8  = cISY_ISEV * y[ISYC];
    const double f9  = cISY_R * y[ISYC];
    const double f10 = cISEV_ICR * y[ISEV];
    const double f11 = cISEV_R * y[ISEV];
    const double f12 = cISEV_D * y[ISEV];
    const double f13 = cICR_D * y[ICR];
    const double f14 = cICR_R * y[ICR];
    dy[S]    = -f0;
    dy[E]    = f0 - f1;
    dy[INS]  = (f1 - f2) - f3;
    dy[INSC] = -f4 - f5;
    dy[ISY]  = (f2 - f6) - f7;
    dy[ISYC] = (f4 - f8) - f9;
    dy[ISEV] = (((f6 + f8) - f10) - f11) - f12;
    dy[ICR]  = (f10 - f13) - f14;
    dy[R]    = ((((f3 + f5) + f7) + f9) + f11) + f14;
    dy[D]    = f12 + f13;
}

// Right-hand side of the node ODE, pop == y (osecir::Model::get_flows + FlowModel::get_derivatives)
template <int G>
__device__ __forceinline__ void rhs(const SegDev& c, Lane<G>& L, const double (&y)[NC], double t, double season,
                                    double (&dy)[NC])
{
    double c_icr, c_d;
    const double kappa = force_of_infection<G>(c, L, y, t, season, c_icr, c_d);
    flows_to_dydt(y, kappa, L.c[LC_E], L.c[LC_INS_ISY], L.c[LC_INS_R], L.c[LC_ISY_ISEV], L.c[LC_ISY_R], c_icr,
                  L.c[LC_ISEV_R], c_d, L.c[LC_ICR_D], L.c[LC_ICR_R], dy);
}

// ------------------------------------------------------------------------------- Cash-Karp 5(4)
// Cash & Karp, ACM TOMS 16 (1990) 201-222; the same ratios odeint's runge_kutta_cash_karp54 uses.
namespace ck
{
constexpr double C2 = 1.0 / 5.0, C3 = 3.0 / 10.0, C4 = 3.0 / 5.0, C5 = 1.0, C6 = 7.0 / 8.0;
constexpr double A21 = 1.0 / 5.0;
constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
constexpr double A41 = 3.0 / 10.0, A42 = -9.0 / 10.0, A43 = 6.0 / 5.0;
constexpr double A51 = -11.0 / 54.0, A52 = 5.0 / 2.0, A53 = -70.0 / 27.0, A54 = 35.0 / 27.0;
constexpr double A61 = 1631.0 / 55296.0, A62 = 175.0 / 512.0, A63 = 575.0 / 13824.0, A64 = 44275.0 / 110592.0,
                 A65 = 253.0 / 4096.0;
constexpr double B1 = 37.0 / 378.0, B3 = 250.0 / 621.0, B4 = 125.0 / 594.0, B6 = 512.0 / 1771.0;
constexpr double E1 = 37.0 / 378.0 - 2825.0 / 27648.0, E3 = 250.0 / 621.0 - 18575.0 / 48384.0,
                 E4 = 125.0 / 594.0 - 13525.0 / 55296.0, E5 = -277.0 / 14336.0, E6 = 512.0 / 1771.0 - 1.0 / 4.0;
} // namespace ck

// One Cash-Karp step from (y, k1 = f(y, t)) with step h: 5th-order solution in yo, embedded error in er.
template <int G>
__device__ __forceinline__ void cash_karp_stages(const SegDev& c, Lane<G>& L, const double (&y)[NC],
                                                 const double (&k1)[NC], double t, double h, double (&yo)[NC],
                                                 double (&er)[NC], bool want_err, double& season_end)
{
    const StageSeasons<G> ss = stage_seasons<G>(L, t, h, ck::C2, ck::C3, ck::C4, ck::C5, ck::C6);
    double k2[NC], k3[NC], k4[NC], k5[NC], yt[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (ck::A21 * h) * k1[i];
    }
    rhs<G>(c, L, yt, t + ck::C2 * h, stage_season<G, 0>(L, ss), k2);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (ck::A31 * h) * k1[i] + (ck::A32 * h) * k2[i];
    }
    rhs<G>(c, L, yt, t + ck::C3 * h, stage_season<G, 1>(L, ss), k3);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (ck::A41 * h) * k1[i] + (ck::A42 * h) * k2[i] + (ck::A43 * h) * k3[i];
    }
    rhs<G>(c, L, yt, t + ck::C4 * h, stage_season<G, 2>(L, ss), k4);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (ck::A51 * h) * k1[i] + (ck::A52 * h) * k2[i] + (ck::A53 * h) * k3[i] + (ck::A54 * h) * k4[i];
    }
    season_end = stage_season<G, 3>(L, ss); // C5 = 1: the season at t + h, the next step's k1 time
    rhs<G>(c, L, yt, t + ck::C5 * h, season_end, k5);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (ck::A61 * h) * k1[i] + (ck::A62 * h) * k2[i] + (ck::A63 * h) * k3[i] + (ck::A64 * h) * k4[i] +
                (ck::A65 * h) * k5[i];
    }
    double k6[NC];
    rhs<G>(c, L, yt, t + ck::C6 * h, stage_season<G, 4>(L, ss), k6);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yo[i] = y[i] + (ck::B1 * h) * k1[i] + (ck::B3 * h) * k3[i] + (ck::B4 * h) * k4[i] + (ck::B6 * h) * k6[i];
    }
    if (want_err) {
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            er[i] = (ck::E1 * h) * k1[i] + (ck::E3 * h) * k3[i] + (ck::E4 * h) * k4[i] + (ck::E5 * h) * k5[i] +
                    (ck::E6 * h) * k6[i];
        }
    }
}

// ------------------------------------------------------------------------------- RKF45 (RKIntegratorCore)
// adapt_rk.h:71-137 default tableau
namespace rkf
{
constexpr double C2 = 0.25, C3 = 3 / 8.0, C4 = 12 / 13.0, C5 = 1.0, C6 = 0.5;
constexpr double A21 = 0.25;
constexpr double A31 = 3 / 32.0, A32 = 9 / 32.0;
constexpr double A41 = 1932 / 2197.0, A42 = -7200 / 2197.0, A43 = 7296 / 2197.0;
constexpr double A51 = 439 / 216.0, A52 = -8.0, A53 = 3680 / 513.0, A54 = -845 / 4104.0;
constexpr double A61 = -8 / 27.0, A62 = 2.0, A63 = -3544 / 2565.0, A64 = 1859 / 4104.0, A65 = -11 / 40.0;
constexpr double BL1 = 25 / 216.0, BL3 = 1408 / 2565.0, BL4 = 2197 / 4104.0, BL5 = -0.2;
constexpr double BH1 = 16 / 135.0, BH3 = 6656 / 12825.0, BH4 = 28561 / 56430.0, BH5 = -9 / 50.0, BH6 = 2 / 55.0;
constexpr double D1 = BH1 - BL1, D3 = BH3 - BL3, D4 = BH4 - BL4, D5 = BH5 - BL5, D6 = BH6 - 0.0;
} // namespace rkf

template <int G>
__device__ __forceinline__ void rkf45_stages(const SegDev& c, Lane<G>& L, const double (&y)[NC],
                                             const double (&k1)[NC], double t, double h, double (&yo)[NC],
                                             double (&er)[NC], double& season_end)
{
    const StageSeasons<G> ss = stage_seasons<G>(L, t, h, rkf::C2, rkf::C3, rkf::C4, rkf::C5, rkf::C6);
    double k2[NC], k3[NC], k4[NC], k5[NC], k6[NC], yt[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (h * rkf::A21) * k1[i];
    }
    rhs<G>(c, L, yt, t + rkf::C2 * h, stage_season<G, 0>(L, ss), k2);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (h * rkf::A31) * k1[i] + (h * rkf::A32) * k2[i];
    }
    rhs<G>(c, L, yt, t + rkf::C3 * h, stage_season<G, 1>(L, ss), k3);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (h * rkf::A41) * k1[i] + (h * rkf::A42) * k2[i] + (h * rkf::A43) * k3[i];
    }
    rhs<G>(c, L, yt, t + rkf::C4 * h, stage_season<G, 2>(L, ss), k4);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (h * rkf::A51) * k1[i] + (h * rkf::A52) * k2[i] + (h * rkf::A53) * k3[i] + (h * rkf::A54) * k4[i];
    }
    season_end = stage_season<G, 3>(L, ss); // C5 = 1: the season at t + h, the next step's k1 time
    rhs<G>(c, L, yt, t + rkf::C5 * h, season_end, k5);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yt[i] = y[i] + (h * rkf::A61) * k1[i] + (h * rkf::A62) * k2[i] + (h * rkf::A63) * k3[i] +
                (h * rkf::A64) * k4[i] + (h * rkf::A65) * k5[i];
    }
    rhs<G>(c, L, yt, t + rkf::C6 * h, stage_season<G, 4>(L, ss), k6);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        // y + dt * (K * b_low) and dt * |K * (b_high - b_low)| (adapt_rk.h:268-271)
        const double sl = k1[i] * rkf::BL1 + k3[i] * rkf::BL3 + k4[i] * rkf::BL4 + k5[i] * rkf::BL5;
        yo[i]           = y[i] + sl * h;
        const double sd = k1[i] * rkf::D1 + k3[i] * rkf::D3 + k4[i] * rkf::D4 + k5[i] * rkf::D5 + k6[i] * rkf::D6;
        er[i]           = h * fabs(sd);
    }
}

template <int G>
__device__ __forceinline__ const SegDev* load_lane(const LaunchDev& ld, Lane<G>& L, double* wx)
{
    L.wx = wx + (threadIdx.x / WAVE) * WAVE;
    L.ws = wx + BLOCK + (threadIdx.x / WAVE) * WAVE;
    constexpr int NPW = WAVE / G;
    const int lb      = xcd_block();
    if (lb >= ld.num_blocks) {
        return nullptr;
    }
    const int seg = ld.block_seg[lb];
    if (seg < 0) {
        return nullptr;
    }
    const SegDev* c = ld.segs + seg;
    const int wave  = (lb - ld.block_first[seg]) * (BLOCK / WAVE) + int(threadIdx.x) / WAVE;
    const int lane  = int(threadIdx.x) % WAVE;
    const int grp   = lane / G;
    L.age           = lane - grp * G;
    L.node          = wave * NPW + grp;
    L.base          = grp * G; // lane index (within the wavefront) of the group's first lane
    if (grp >= NPW || L.node >= c->num_nodes) {
        return nullptr;
    }
    if (c->node_order) {
        L.node = c->node_order[L.node]; // same graph node of consecutive runs in one wavefront
    }
    L.np            = c->nodep[L.node];
    const double* p = c->lcoef + (size_t(L.node) * G + L.age) * NUM_LC;
#pragma unroll
    for (int k = 0; k < NUM_LC; ++k) {
        L.c[k] = p[k];
    }
    const double* B = c->cbase + (size_t(L.node) * G + L.age) * G;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        L.crow[j] = B[j];
    }
    return c;
}

__device__ __forceinline__ double detect_nondetection(const NodeP& np, double t)
{
    // test_commuters (model.h:717-742): nondetection applies in [start, end)
    return (t >= np.det_start && t < np.det_end) ? np.nondetect : 1.0;
}

// mobility factor of the ISy compartment of the lane's age group (model.h:687-715)
template <int G>
__device__ __forceinline__ double mobility_factor(const Lane<G>& L, const double (&y)[NC])
{
    if (!(L.np.flags & NF_TTC_ACTIVE)) {
        return L.c[LC_RISK];
    }
    const double ttr = group_sum<G>(L.c[LC_TTR] * y[INS], L.base);
    return smoother_cosine(ttr, L.np.ttc, L.np.ttc_hi, L.c[LC_RISK], L.c[LC_MAXRISK]);
}

// ------------------------------------------------------------------------------- clock
// One GraphSimulation::advance loop iteration per instance (graph_simulation.h:101-120) and the
// departure/return alternation of MobilityEdge::apply_mobility (metapopulation_mobility_instant.h:544-623).
__device__ __forceinline__ void clock_one(InstDev& I)
{
    I.c_active  = 0;
    I.c_action  = 0;
    I.c_out_day = -1;
    I.fx_pass   = 0;
    I.fx_state  = 0;
    I.fx_changed = 0;
    I.fx_count   = 0;
    if (I.status != 0 || !(I.t < I.tmax)) {
        return;
    }
    double step = I.dt_graph;
    if (I.t + step > I.tmax) {
        step = I.tmax - I.t;
    }
    // single-model runs end exactly at tmax (Simulation::advance(tmax)); graph runs at t + dt like
    // SimulationNode::advance(t, dt)
    const double tn = (I.exact_end && step != I.dt_graph) ? I.tmax : I.t + step;
    I.c_active      = 1;
    I.c_t0          = I.t;
    I.c_t1          = tn;
    I.c_step        = step;
    if (I.estruct >= 0) {
        if (I.pending && I.ret_time <= tn) {
            if (I.toggle) {
                I.c_action = 2;
                I.c_tdep   = I.t_dep;
                I.pending  = 0;
            }
            else {
                I.status = MGO_ERR_UNSUPPORTED; // return and departure in one call: more than one pending move
            }
        }
        else if (!I.toggle) {
            if (I.pending) {
                I.status = MGO_ERR_UNSUPPORTED;
            }
            else {
                I.c_action = 1;
                I.pending  = 1;
                I.t_dep    = tn;
                I.ret_time = tn + step;
            }
        }
        I.toggle ^= 1;
    }
    if (I.c_action == 2) {
        I.fx_epoch += FX_EPOCH_STEP; // fix-up records of earlier return phases read as unfixed
    }
    I.t             = tn;
    const double dn = rint(tn);
    if (fabs(tn - dn) <= 1e-10) {
        const int d = int(dn) - I.day0;
        if (d >= 0 && d < I.ndays) {
            I.c_out_day = d;
        }
    }
}

// One block: the first threads clear the cycle flags, then every instance advances its clock and flags the
// phases it needs (every writer stores 1).
__global__ void __launch_bounds__(1024) clock_kernel(InstDev* inst, int n, int32_t* cyc)
{
    if (cyc) {
        if (threadIdx.x < CYC_FLAGS) {
            cyc[threadIdx.x] = 0;
        }
        __syncthreads();
    }
    for (int i = int(threadIdx.x); i < n; i += int(blockDim.x)) {
        InstDev& I = inst[i];
        clock_one(I);
        if (cyc && I.c_action == 1) {
            cyc[CYC_DEPART] = 1;
        }
        if (cyc && I.c_action == 2) {
            cyc[CYC_RETURN] = 1;
        }
    }
}

// daily output of local node v: [day][v][g][10]
__device__ __forceinline__ double* out_ptr(const InstDev& I, int v, int age, int G, int day)
{
    return I.out + (size_t(day) * I.num_nodes + v) * size_t(G * NC) + size_t(age) * NC;
}

// linear_interpolation (math/interpolation.h:44-48): y_a + w (y_b - y_a), rounded as the reference (no FMA)
__device__ __forceinline__ double lerp_ref(double ya, double yb, double w)
{
#pragma clang fp contract(off)
    return ya + w * (yb - ya);
}

// Daily value of the output day d = c_out_day at the end of a phase, from the lane's state y after the phase's
// edge action (apply_mobility changes the last value of the node's TimeSeries in place), as
// interpolate_simulation_result (analyze_result.h:235-261) takes it from the node's time points: between the
// state before the node's last step and y when that step ends at or after d (weight 1 when it ends on d);
// else y, the last value so far: then the next phase's first step encloses d and overwrites the day.
template <int G>
__device__ __forceinline__ void write_phase_end_day(const SegDev& c, const InstDev& I, size_t lane, int v, int age,
                                                    const double (&y)[NC])
{
    const double* pr = c.yint + lane * YINT_STRIDE;
    const double w   = pr[YINT_W];
    double* op       = out_ptr(I, v, age, G, I.c_out_day);
    if (w >= 0.0) {
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            op[i] = lerp_ref(pr[i], y[i], w);
        }
    }
    else {
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            op[i] = y[i];
        }
    }
}

// segment node of an instance's local node v
__device__ __forceinline__ size_t seg_node(const InstDev& I, int v)
{
    return size_t(I.node_base) + size_t(v) * size_t(I.node_stride);
}

// Return coefficients of the lane's node as a destination (calculate_mobility_returns,
// metapopulation_mobility_instant.h:460-471): rates and force of infection of the state pop at the departure
// time, frozen for the Euler step of the returning travelers.
template <int G>
__device__ __forceinline__ void return_coefficients(const SegDev& c, Lane<G>& L, const InstDev& I,
                                                    const double (&pop)[NC])
{
    double c_icr, c_d;
    const double kappa = force_of_infection<G>(c, L, pop, I.c_tdep, season_at<G>(L, I.c_tdep), c_icr, c_d);
    double* rc         = c.rc + (size_t(L.node) * G + L.age) * RC_STRIDE;
    rc[RC_KAPPA]       = kappa;
    rc[RC_E]           = L.c[LC_E];
    rc[RC_INS_ISY]     = L.c[LC_INS_ISY];
    rc[RC_INS_R]       = L.c[LC_INS_R];
    rc[RC_ISY_ISEV]    = L.c[LC_ISY_ISEV];
    rc[RC_ISY_R]       = L.c[LC_ISY_R];
    rc[RC_ISEV_ICR]    = c_icr;
    rc[RC_ISEV_R]      = L.c[LC_ISEV_R];
    rc[RC_ISEV_D]      = c_d;
    rc[RC_ICR_D]       = L.c[LC_ICR_D];
    rc[RC_ICR_R]       = L.c[LC_ICR_R];
}

// ------------------------------------------------------------------------------- dynamic NPIs
// The check at the start of one piece of osecir::Simulation::advance (models/ode_secir/model.h:301-349) for
// the lane's node at time t with state y, and the length of the piece (model.h:351-357: at most one day,
// ending at the active NPI's end). Every lane of the node reaches the same decision; the first lane of the
// group implements the NPI on the node's damping list and cache (mgo_npi.h). Returns true in `changed`
// when the contact matrix changed.
template <int G>
__device__ double npi_piece(const SegDev& c, Lane<G>& L, InstDev& I, const double (&y)[NC], double t, double tmax,
                            bool& changed)
{
    NpiNode* nn         = c.npi + L.node;
    const double sum    = group_sum<G>(y[ISY], L.base); // InfectedSymptoms summed over the groups in order
    const double begin  = nn->begin, end = nn->end;
    double t_end        = nn->t_end;
    changed             = false;
    // floating_point_greater_equal(a, b, 1e-10) (math/floating_point.h:96-100)
    auto ge = [](double a, double b) {
        return !(a - b < -(1e-10 + DBL_MIN * fmax(fabs(a), fabs(b))));
    };
    if (ge(t, begin) && t < end) {
        const double inf_rel = sum / nn->total * nn->base;
        int k                = 0;
        while (k < nn->thr_count && !(inf_rel > c.npi_thr[nn->thr_first + k])) {
            ++k;
        }
        const bool expired = ge(t, t_end);
        if (k < nn->thr_count) {
            const double thr   = c.npi_thr[nn->thr_first + k];
            const double value = nn->value;
            if (thr > value || expired) {
                const bool renewal = expired && !(thr > value);
                const double delay = renewal ? 0.0 : (t > 0 ? nn->delay : 0.0);
                if (t + delay < end) {
                    const double t_start = t + delay;
                    const double te      = (t_start + nn->duration < end) ? t_start + nn->duration : end;
                    const double tsd     = (!renewal && t_start > 0.0) ? t_start + 1.0 : t_start;
                    t_end                = te;
                    changed              = true;
                    if (L.age == 0) {
                        nn->value = thr;
                        nn->t_end = te;
                        const int gi = nn->thr_first + k;
                        const int d0 = c.npi_thr_damp[gi], d1 = c.npi_thr_damp[gi + 1];
                        int count    = nn->list_count;
                        bool ok      = npi_implement(nn->list, count, nn->list_cap, G, true,
                                                     c.npi_damp + d0 * (3 + G), d1 - d0, tsd, te + 1.0, nn->work);
                        const int m  = ok ? npi_build_cache(nn->list, count, G * G, nn->list_cap, nn->cache_t,
                                                            nn->cache_m, nn->work)
                                          : -1;
                        if (m < 0) {
                            I.status = MGO_ERR_OUT_OF_RANGE; // damping list or level capacity exceeded
                        }
                        else {
                            nn->list_count  = count;
                            nn->cache_count = m;
                            const_cast<NodeP*>(c.nodep)[L.node].damp_cnt = m;
                        }
                    }
                }
            }
        }
    }
    if (changed) {
        __threadfence(); // the group's other lanes read the new cache written by its first lane
        L.np.damp_cnt = __shfl(*static_cast<volatile int32_t*>(&nn->cache_count), L.base);
    }
    double dt_eff = (tmax - t < 1.0) ? tmax - t : 1.0;
    if (t < t_end) {
        dt_eff = (t_end - t < dt_eff) ? t_end - t : dt_eff;
    }
    return dt_eff;
}

// Out-edge dampings of the lane's node at the cycle's edge action (MobilityEdge::apply_mobility at t = c_t1,
// metapopulation_mobility_instant.h:549-566): the edge NPI check on the node's last state (no delay, no
// directive; begin t, end t + duration, per-group mobility dampings), then on a departure the damping factor
// D of the lane's age group at t (Dampings::get_matrix_at, damping.h:381-396) for the edge kernels.
template <int G>
__device__ void out_damping(const SegDev& c, Lane<G>& L, InstDev& I, const double (&y)[NC])
{
    NpiNode* od    = c.out_npi + L.node;
    const double t = I.c_t1;
    if (od->thr_count > 0) {
        const double sum     = group_sum<G>(y[ISY], L.base);
        const double inf_rel = sum / od->total * od->base;
        int k                = 0;
        while (k < od->thr_count && !(inf_rel > c.out_thr[od->thr_first + k])) {
            ++k;
        }
        const double te    = od->t_end;
        const bool expired = !(t - te < -(1e-10 + DBL_MIN * fmax(fabs(t), fabs(te))));
        bool changed       = false;
        if (k < od->thr_count) {
            const double thr = c.out_thr[od->thr_first + k];
            if (thr > od->value || expired) {
                changed = true;
                if (L.age == 0) {
                    const double t_end = t + od->duration;
                    od->value          = thr;
                    od->t_end          = t_end;
                    const int gi       = od->thr_first + k;
                    const int d0 = c.out_thr_damp[gi], d1 = c.out_thr_damp[gi + 1];
                    int count    = od->list_count;
                    const bool ok = npi_implement(od->list, count, od->list_cap, G, false, c.out_damp + d0 * (3 + G),
                                                  d1 - d0, t, t_end, od->work);
                    const int m   = ok ? npi_build_cache(od->list, count, G, od->list_cap, od->cache_t, od->cache_m,
                                                         od->work)
                                       : -1;
                    if (m < 0) {
                        I.status = MGO_ERR_OUT_OF_RANGE; // damping list or level capacity exceeded
                    }
                    else {
                        od->list_count  = count;
                        od->cache_count = m;
                    }
                }
            }
        }
        if (changed) {
            __threadfence(); // the group's other lanes read the cache written by its first lane
        }
    }
    if (I.c_action == 1) {
        const int cnt    = *static_cast<volatile int32_t*>(&od->cache_count);
        const double* tk = od->cache_t;
        int k            = 1;
        while (k < cnt - 1 && !(tk[k] > t)) {
            ++k;
        }
        const double xr = tk[k], xl = xr - 1.0;
        const double Dl = od->cache_m[(k - 1) * G + L.age], Dr = od->cache_m[k * G + L.age];
        double d;
        if (t <= xl) {
            d = Dl;
        }
        else if (t >= xr) {
            d = Dr;
        }
        else {
            d = 0.5 * (Dl - Dr) * cos(MGO_PI / (xr - xl) * (t - xl)) + 0.5 * (Dl + Dr);
        }
        c.out_d[size_t(L.node) * G + L.age] = d;
    }
}

// ------------------------------------------------------------------------------- node phase
// SystemIntegrator::advance (integrator.h:184-249) of every node to the cycle's end time with the
// integrator core K: odeint Cash-Karp controller (stepper_wrapper.h:112-163), RKIntegratorCore
// (adapt_rk.h:217-298), fixed Cash-Karp (stepper_wrapper.h:240-255) or Euler (euler.h:54-62).
// Loop is flattened to one try_step attempt per iteration so that a rejected node only idles through
// the k1 evaluation of its neighbours in the wavefront.
constexpr uint32_t kMaxAttemptsPerPhase = 1u << 20;

template <int G, int K, bool NPI = false>
__device__ __forceinline__ void node_body(const LaunchDev& ld, InstDev* __restrict__ inst, double* wx)
{
    Lane<G> L;
    const SegDev* cp = load_lane<G>(ld, L, wx);
    if (!cp) {
        return;
    }
    const SegDev& c = *cp;
    InstDev& I = inst[L.np.inst];
    if (!I.c_active) {
        return;
    }
    double* yp = c.y + (size_t(L.node) * G + L.age) * NC;
    double y[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        y[i] = yp[i];
    }
    double t            = c.tnode[L.node];
    double dt           = c.dtnode[L.node];
    const double t_phase = I.c_t1;
    // Without dynamic NPIs the phase is one SystemIntegrator::advance to c_t1; with them it is the pieces of
    // osecir::Simulation::advance (at most a day each, ending at NPI ends), each starting with the NPI check
    const bool npi      = NPI && (L.np.flags & NF_NPI) != 0; // NPI: instantiation for classes with NPIs
    double t_loop       = t; // osecir::Simulation::advance's loop time (model.h:301)
    double tmax         = t_phase;
    double t0           = t;
    const double dt_max = I.dt_max;
    double dt_min       = I.dt_min;
    double dt_restore   = 0.0;
    const double eps    = 1e-12;
    // output days strictly inside this phase are interpolated here; a day at the phase end is written
    // after the edge action
    // (the first is the day after the node's own time: a fixed-step node can end a phase just below the day
    // the phase ended on, and then its first step here encloses that day)
    const int ndays  = I.ndays;
    int next_day     = max(0, int(floor(t)) + 1 - I.day0);
    const double tde = I.c_t1 - 1e-10;

    bool active  = false;
    bool need_k1 = true;
    // season of the next k1: an accepted step ends at t + h, where its fifth stage evaluated the season already
    bool have_season   = false;
    double season_next = 0.0, season_end = 0.0;
    bool fail = true, valid = true, first = true, stalled = false;
    uint32_t nsteps = 0, natt = 0;
    double k1[NC];
    for (;;) { // pieces of the node advance (one without dynamic NPIs)
        bool changed = false;
        if (npi) {
            tmax = t_loop + npi_piece<G>(c, L, I, y, t_loop, t_phase, changed);
        }
        if (I.c_action == 2 && (first || changed)) {
            // return coefficients of this node as a destination: its state after the departures, at the
            // departure time (the `pop` of calculate_mobility_returns, metapopulation_mobility_instant.h:460-471),
            // with the contact matrix as it is after the node's NPI checks of this phase
            double pop[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                pop[i] = first ? y[i] : c.npi[L.node].y0[L.age * NC + i];
                if (first && npi) {
                    c.npi[L.node].y0[L.age * NC + i] = y[i];
                }
            }
            return_coefficients<G>(c, L, I, pop);
        }
        first  = false;
        t0     = t;
        active = fabs((tmax - t) / (tmax - t0)) > eps;
        while (active) {
            if (natt >= kMaxAttemptsPerPhase) {
                // a step size driven below any progress (dt_min ~ DBL_MIN with a persistent error): stop this
                // node and report instead of spinning; the reference would not terminate here
                if (L.age == 0) {
                    I.status = MGO_ERR_OUT_OF_RANGE;
                }
                stalled = true;
                break;
            }
            if (need_k1) {
                if (dt > tmax - t) {
                    dt_restore = dt;
                    dt         = tmax - t;
                    dt_min     = fmin(tmax - t, dt_min);
                }
                rhs<G>(c, L, y, t, have_season ? season_next : season_at<G>(L, t), k1);
                have_season = false;
                need_k1 = false;
                fail    = true;
                valid   = true;
                if (K == MGO_INTEGRATOR_RKF45) {
                    dt = fmin(dt, dt_max);
                }
            }
            double yo[NC], er[NC];
            bool done       = false;
            const double ta = t;
            ++natt;
            if (K == MGO_INTEGRATOR_CASH_KARP54) {
                if (dt < dt_min) {
                    valid = false;
                    dt    = dt_min;
                }
                if (!(dt - dt_max <= DBL_EPSILON)) {
                    dt   = dt_max; // check_step_size_limit: rejected without a step
                    fail = true;
                }
                else {
                    cash_karp_stages<G>(c, L, y, k1, t, dt, yo, er, true, season_end);
                    // max_i |err_i| / (abs + rel (|y_i| + dt |k1_i|)) (odeint default_error_checker): the index of
                    // the largest quotient by cross-multiplied comparison, then one division
                    double en = 0.0, ed = 1.0;
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        const double n = fabs(er[i]);
                        const double d = I.abs_tol + I.rel_tol * (fabs(y[i]) + dt * fabs(k1[i]));
                        if (n * ed > en * d) {
                            en = n;
                            ed = d;
                        }
                    }
                    double e = en / ed;
                    // maximum over the node's lanes through the wavefront's slots (the same fmax order)
                    L.wx[threadIdx.x % WAVE] = e;
                    __builtin_amdgcn_wave_barrier();
                    e = L.wx[L.base];
#pragma unroll
                    for (int j = 1; j < G; ++j) {
                        e = fmax(e, L.wx[L.base + j]);
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (e > 1.0) {
                        dt *= fmax(0.9 * step_pow(e, -1.0 / 3.0), 0.2);
                        dt   = fmin(dt, dt_max);
                        fail = true;
                    }
                    else {
                        t += dt;
                        have_season = true;
                        season_next = season_end;
                        if (e < 0.5) {
                            e = fmax(1.0 / 3125.0, e);
                            dt *= 0.9 * step_pow(e, -0.2);
                            dt = fmin(dt, dt_max);
                        }
                        fail = false;
                    }
                }
                done = !(fail && valid);
                if (done) {
                    dt = fmax(dt, dt_min);
                    if (fail) {
                        t += dt_min;
                    }
                }
            }
            else if (K == MGO_INTEGRATOR_RKF45) {
                if (dt < dt_min) {
                    valid = false;
                    dt    = dt_min;
                }
                rkf45_stages<G>(c, L, y, k1, t, dt, yo, er, season_end);
                bool conv = true;
                double q  = INFINITY;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    const double ep = I.abs_tol + fabs(yo[i]) * I.rel_tol;
                    conv            = conv && (er[i] <= ep);
                    q               = fmin(q, ep / er[i]);
                }
                conv = group_all<G>(conv, L.base);
                q    = group_min<G>(q, L.base);
                if (conv || !valid) {
                    t += dt;
                    have_season = true;
                    season_next = season_end;
                }
                double dt_new = dt * step_pow(q, 1.0 / 5.0);
                dt_new *= 0.9;
                dt   = fmin(dt_new, dt_max);
                done = conv || !valid;
                if (done) {
                    dt = fmax(dt, dt_min);
                }
            }
            else if (K == MGO_INTEGRATOR_CASH_KARP54_FIXED) {
                cash_karp_stages<G>(c, L, y, k1, t, dt, yo, er, false, season_end);
                t += dt;
                have_season = true;
                season_next = season_end;
                done = true;
            }
            else { // Euler
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    yo[i] = y[i] + dt * k1[i];
                }
                t += dt;
                done = true;
            }
            if (done) {
                ++nsteps;
                // interpolate_simulation_result (analyze_result.h:235-252): days in (ta, t] strictly inside
                // the phase
                while (next_day < ndays && double(next_day + I.day0) <= t && double(next_day + I.day0) < tde) {
                    const double dd = double(next_day + I.day0);
                    if (dd > ta) {
                        const double w = (dd - ta) / (t - ta);
                        double* op     = out_ptr(I, L.np.local, L.age, G, next_day);
#pragma unroll
                        for (int i = 0; i < NC; ++i) {
                            op[i] = lerp_ref(y[i], yo[i], w);
                        }
                    }
                    ++next_day;
                }
                active = fabs((tmax - t) / (tmax - t0)) > eps;
                if (!active && I.c_out_day >= 0) {
                    // the last step of the phase (so far): the phase-end day is interpolated from it after the
                    // edge action (write_phase_end_day)
                    const double dd = double(I.c_out_day + I.day0);
                    double* pr      = c.yint + (size_t(L.node) * G + L.age) * YINT_STRIDE;
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        pr[i] = y[i];
                    }
                    pr[YINT_W] = (t >= dd) ? (dd - ta) / (t - ta) : -1.0;
                }
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    y[i] = yo[i];
                }
                need_k1 = true;
            }
        }
        if (!npi || stalled) {
            break;
        }
        // end of a piece: SystemIntegrator::advance's exit (integrator.h:244-248), then the next piece
        dt         = fmax(dt, dt_restore);
        dt_restore = 0.0;
        dt_min     = I.dt_min;
        t_loop     = tmax;
        if (!(t_loop < t_phase)) {
            break;
        }
        need_k1 = true;
    }
    dt = fmax(dt, dt_restore);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yp[i] = y[i];
    }
    if (L.age == 0) {
        c.tnode[L.node]  = t;
        c.dtnode[L.node] = dt;
        c.steps[L.node] += nsteps;
        c.attempts[L.node] += natt;
    }
    const size_t off = (size_t(L.node) * G + L.age) * NC;
    if (NPI && (L.np.flags & NF_OUT_DAMP) && I.c_action != 0) {
        out_damping<G>(c, L, I, y);
    }
    if (I.c_action == 1) {
        // departure row: SimulationNode::m_last_state, the ISy mobility factor and the commuter
        // nondetection factor at the departure time
        double* row = c.snap + (size_t(L.node) * G + L.age) * SNAP_STRIDE;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            row[i] = y[i];
        }
        row[SNAP_FAC] = mobility_factor<G>(L, y);
        row[SNAP_ND]  = (L.np.flags & NF_DETECTION) ? detect_nondetection(L.np, I.c_t1) : 1.0;
    }
    else if (I.c_action == 2) {
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            c.ypre[off + i] = y[i];
        }
    }
    else if (I.c_out_day >= 0) {
        write_phase_end_day<G>(c, I, size_t(L.node) * G + L.age, L.np.local, L.age, y);
    }
}

// ------------------------------------------------------------------------------- edge phase
// Lane -> (segment, node, age group) without loading the lane's coefficients (edge kernels need none).
template <int G>
__device__ __forceinline__ const SegDev* lane_map(const LaunchDev& ld, int& node, int& age, int& base)
{
    constexpr int NPW = WAVE / G;
    const int lb      = xcd_block();
    if (lb >= ld.num_blocks) {
        return nullptr;
    }
    const int seg = ld.block_seg[lb];
    if (seg < 0) {
        return nullptr;
    }
    const SegDev* c = ld.segs + seg;
    const int wave  = (lb - ld.block_first[seg]) * (BLOCK / WAVE) + int(threadIdx.x) / WAVE;
    const int lane  = int(threadIdx.x) % WAVE;
    const int grp   = lane / G;
    age             = lane - grp * G;
    node            = wave * NPW + grp;
    base            = grp * G;
    if (grp >= NPW || node >= c->num_nodes) {
        return nullptr;
    }
    return c;
}

// departure row of lane slot `lane` (16-byte loads)
__device__ __forceinline__ void load_row(const double* snap, size_t lane, double (&r)[SNAP_STRIDE])
{
    const double2* p = reinterpret_cast<const double2*>(snap + lane * SNAP_STRIDE);
#pragma unroll
    for (int i = 0; i < SNAP_STRIDE / 2; ++i) {
        const double2 v = p[i];
        r[2 * i]        = v.x;
        r[2 * i + 1]    = v.y;
    }
}

// traveler vector of one edge from its source's departure row (metapopulation_mobility_instant.h:608-615:
// snapshot * coefficients * mobility factors, then test_commuters' nondetection on INS and ISy if with_nd).
// The coefficients at the departure time are B - D (B - M) (DampingMatrixExpression::get_matrix_at,
// contact_matrix.h:190-194) with the source's out-edge damping factor D; mn: the entry's uniform minimum
// (entry index k of the in- or out-entries), or NULL.
__device__ __forceinline__ void traveler_vector(const EdgeStructDev& es, const EdgeEnt& e, const double* mn, int k,
                                                double D, const double (&row)[SNAP_STRIDE], int G, int age,
                                                bool with_nd, double (&m)[NC])
{
    if (e.cidx < 0) {
        double cv = e.cval;
        if (D != 0.0) {
            cv = e.cval - D * (e.cval - (mn ? mn[k] : 0.0));
        }
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            m[i] = row[i] * cv;
        }
    }
    else {
        const size_t off = size_t(e.cidx) * G * NC + size_t(age) * NC;
        const double* cv = es.cvec + off;
        if (D != 0.0) {
            const double* mv = es.cvec_min ? es.cvec_min + off : nullptr;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                m[i] = row[i] * (cv[i] - D * (cv[i] - (mv ? mv[i] : 0.0)));
            }
        }
        else {
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                m[i] = row[i] * cv[i];
            }
        }
    }
    m[ISY] = m[ISY] * row[SNAP_FAC];
    if (with_nd && row[SNAP_ND] != 1.0) {
        m[ISY] *= row[SNAP_ND];
        m[INS] *= row[SNAP_ND];
    }
}

// one explicit Euler step of the travelers with the destination's return coefficients
// (calculate_mobility_returns, metapopulation_mobility_instant.h:460-471; euler.h:54-62)
__device__ __forceinline__ void return_euler(const double* rc, double h, double (&m)[NC])
{
    double dy[NC];
    flows_to_dydt(m, rc[RC_KAPPA], rc[RC_E], rc[RC_INS_ISY], rc[RC_INS_R], rc[RC_ISY_ISEV], rc[RC_ISY_R],
                  rc[RC_ISEV_ICR], rc[RC_ISEV_R], rc[RC_ISEV_D], rc[RC_ICR_D], rc[RC_ICR_R], dy);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        m[i] = m[i] + h * dy[i];
    }
}

template <int G>
__device__ __forceinline__ bool group_any(bool v, int base)
{
    return ((__ballot(v) >> base) & ((1ull << G) - 1)) != 0;
}

// MobilityEdge::apply_mobility return branch fix-up (metapopulation_mobility_instant.h:576-596): every
// compartment of the age group that would go negative at the destination is taken from the group's
// largest remaining compartment instead (remaining computed once, first maximum wins).
__device__ __forceinline__ bool fixup_negative(const double (&y)[NC], double (&m)[NC])
{
    double rem[NC];
    bool neg = false;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        rem[i] = y[i] - m[i];
        neg    = neg || (rem[i] < 0);
    }
    if (neg) {
        int mx = 0;
#pragma unroll
        for (int q = 1; q < NC; ++q) {
            if (rem[q] > rem[mx]) {
                mx = q;
            }
        }
        for (int j = 0; j < NC; ++j) {
            if (rem[j] < 0) {
                m[mx] -= rem[j];
                m[j] += rem[j];
            }
        }
    }
    return neg;
}

// Departures at t_dep = c_t1 (metapopulation_mobility_instant.h:605-621): every edge (s, d) adds its
// traveler vector to d and subtracts it from s, edges in the global (start, end) order. Node v sees its
// edges as in[s < v], out[all], in[s > v]; the lane keeps its age group's compartments in registers and
// walks that sequence (a deterministic segmented sum in the reference's order, no atomics).
template <int G>
__device__ __forceinline__ void depart_body(const LaunchDev& ld, InstDev* __restrict__ inst)
{
    if (ld.cyc && !ld.cyc[CYC_DEPART]) {
        return; // nobody departs in this cycle
    }
    int node, age, base;
    const SegDev* cp = lane_map<G>(ld, node, age, base);
    if (!cp) {
        return;
    }
    const SegDev& c = *cp;
    InstDev& I      = inst[c.nodep[node].inst];
    if (!I.c_active || I.c_action != 1) {
        return;
    }
    const EdgeStructDev& es = c.estructs[I.estruct];
    const int v             = c.nodep[node].local;
    const int in_b = es.in_ptr[v], in_e = es.in_ptr[v + 1];
    const int in_s          = in_b + es.split[v];
    const int out_b = es.out_ptr[v], out_e = es.out_ptr[v + 1];
    const EdgeEnt* ine      = es.in_ent;
    const EdgeEnt* oute     = es.out_ent;
    const size_t lane_off   = size_t(node) * G + age;
    const size_t src_base   = size_t(I.node_base) * G + age;   // lane of local node 0
    const size_t src_step   = size_t(I.node_stride) * G;       // lanes between local nodes
    double* yp              = c.y + lane_off * NC;
    double y[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        y[i] = yp[i];
    }
    auto in_range = [&](int k, int end) {
        for (; k < end; ++k) {
            const EdgeEnt e = ine[k];
            double r[SNAP_STRIDE];
            const size_t sl = src_base + size_t(e.nbr) * src_step;
            load_row(c.snap, sl, r);
            double m[NC];
            traveler_vector(es, e, es.in_min, k, c.out_d ? c.out_d[sl] : 0.0, r, G, age, true, m);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                y[i] += m[i];
            }
        }
    };
    in_range(in_b, in_s);
    if (out_b < out_e) {
        double sv[SNAP_STRIDE];
        load_row(c.snap, lane_off, sv);
        const double nd_v = sv[SNAP_ND];
        const double D_v  = c.out_d ? c.out_d[lane_off] : 0.0;
        for (int k = out_b; k < out_e; ++k) {
            double m[NC];
            traveler_vector(es, oute[k], es.out_min, k, D_v, sv, G, age, false, m);
            if (nd_v != 1.0) {
                // test_commuters on the source: detected commuters stay home as confirmed cases
                y[INS] -= m[INS] * (1 - nd_v);
                y[INSC] += m[INS] * (1 - nd_v);
                y[ISY] -= m[ISY] * (1 - nd_v);
                y[ISYC] += m[ISY] * (1 - nd_v);
                m[ISY] *= nd_v;
                m[INS] *= nd_v;
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                y[i] -= m[i];
            }
        }
    }
    in_range(in_s, in_e);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yp[i] = y[i];
    }
    if (I.c_out_day >= 0) {
        write_phase_end_day<G>(c, I, lane_off, v, age, y);
    }
}

// newest fix-up record version of an in-edge written in [lo, hi] (slot, -1 = unfixed or none)
__device__ __forceinline__ int32_t record_slot(const FxRec* r, int32_t lo, int32_t hi)
{
    const FxRec a = r[0];
    const FxRec b = r[1];
    const bool va = a.tag >= lo && a.tag <= hi;
    const bool vb = b.tag >= lo && b.tag <= hi;
    if (va && (!vb || a.tag > b.tag)) {
        return a.slot;
    }
    return vb ? b.slot : -1;
}

// Return walk of one node (returns at c_t1 of the travelers that left at c_tdep,
// metapopulation_mobility_instant.h:568-603) from the pre-return state, visiting the node's edges in the
// reference's global (start, end) order.
//
// The fix-up of edge (s, d) depends on d's state at that point of the sequence, which includes d's
// earlier out-edges (d, x) with their own fix-ups decided at x. Pass p applies the fix-ups of the node's
// in-edges locally and takes the out-edge vectors from the records as of pass p-1 (pass 0: unfixed). The
// dependencies follow the global edge order (a DAG), so the records reach the reference's sequential
// values after at most (longest dependency chain + 1) passes; a pass in which no record changes
// reproduces the sequential result exactly. A node whose out-edge records did not change in pass p-1
// would repeat its walk of pass p-1 unchanged, so pass p only re-walks the nodes marked dirty then.
template <int G, bool FIRST>
__device__ __forceinline__ void return_body(const LaunchDev& ld, InstDev* __restrict__ inst, int launch_pass)
{
    if (ld.cyc && !ld.cyc[FIRST ? CYC_RETURN : CYC_PASS0 + launch_pass]) {
        return; // nobody returns in this cycle, or no instance needs this fix-up pass
    }
    int node, age, base;
    const SegDev* cp = lane_map<G>(ld, node, age, base);
    if (!cp) {
        return;
    }
    const SegDev& c = *cp;
    InstDev& I      = inst[c.nodep[node].inst];
    if (!I.c_active || I.c_action != 2) {
        return;
    }
    const int pass = FIRST ? 0 : I.fx_pass;
    if (FIRST ? (I.fx_state != 0) : (I.fx_state != 1 || pass == 0)) {
        return;
    }
    const int32_t tag = I.fx_epoch + pass;
    // marks of pass p-1 are in plane (p-1)&1; this pass writes plane p&1
    int32_t* dirty_out = c.fx_dirty + size_t(pass & 1) * c.num_nodes;
    if (!FIRST && c.fx_dirty[size_t((pass - 1) & 1) * c.num_nodes + node] != tag - 1) {
        return;
    }
    const EdgeStructDev& es = c.estructs[I.estruct];
    const int v             = c.nodep[node].local;
    const int in_b = es.in_ptr[v], in_e = es.in_ptr[v + 1];
    const int in_s          = in_b + es.split[v];
    const int out_b = es.out_ptr[v], out_e = es.out_ptr[v + 1];
    const EdgeEnt* ine      = es.in_ent;
    const EdgeEnt* oute     = es.out_ent;
    const size_t lane_off   = size_t(node) * G + age;
    const size_t src_base   = size_t(I.node_base) * G + age;   // lane of local node 0
    const size_t src_step   = size_t(I.node_stride) * G;       // lanes between local nodes
    FxRec* recs             = c.fx_rec + 2 * I.fx_base;
    double* pool            = c.fx_pool + I.fx_pool_base;
    const int32_t lo        = I.fx_epoch;
    const double h          = I.c_step;
    const double* rc_v      = c.rc + lane_off * RC_STRIDE;

    double y[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        y[i] = c.ypre[lane_off * NC + i];
    }
    bool changed = false;
    auto in_edge = [&](int k) {
        const EdgeEnt e = ine[k];
        double row[SNAP_STRIDE];
        const size_t sl = src_base + size_t(e.nbr) * src_step;
        load_row(c.snap, sl, row);
        double m[NC];
        traveler_vector(es, e, es.in_min, k, c.out_d ? c.out_d[sl] : 0.0, row, G, age, true, m);
        return_euler(rc_v, h, m);
        const bool any = group_any<G>(fixup_negative(y, m), base);
        const int32_t ps = FIRST ? -1 : record_slot(recs + 2 * k, lo, tag - 1);
        bool diff        = (ps >= 0) != any;
        if (any && !diff) {
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                diff = diff || (pool[(size_t(ps) * G + age) * NC + i] != m[i]);
            }
        }
        diff = group_any<G>(diff, base);
        if (diff) {
            int32_t slot = -1;
            if (any) {
                if (age == 0) {
                    slot = atomicAdd(&I.fx_count, 1);
                }
                slot = __shfl(slot, base);
                if (slot < FX_SLOTS) {
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        pool[(size_t(slot) * G + age) * NC + i] = m[i];
                    }
                }
                else {
                    slot = -1;
                    if (age == 0) {
                        I.fx_state = 2; // record pool full: resolve this phase with the sequential replay
                    }
                }
            }
            if (age == 0) {
                FxRec* r        = recs + 2 * k;
                const int older = (r[0].tag <= r[1].tag) ? 0 : 1;
                r[older]        = FxRec{tag, slot};
                dirty_out[seg_node(I, e.nbr)] = tag; // the source re-walks in the next pass
            }
            changed = true;
        }
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            y[i] -= m[i];
        }
    };
    for (int k = in_b; k < in_s; ++k) {
        in_edge(k);
    }
    if (out_b < out_e) {
        double sv[SNAP_STRIDE];
        load_row(c.snap, lane_off, sv);
        const double D_v = c.out_d ? c.out_d[lane_off] : 0.0;
        for (int k = out_b; k < out_e; ++k) {
            const EdgeEnt e = oute[k];
            double m[NC];
            const int32_t ps = FIRST ? -1 : record_slot(recs + 2 * e.peer, lo, tag - 1);
            if (ps >= 0) {
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    m[i] = pool[(size_t(ps) * G + age) * NC + i];
                }
            }
            else {
                traveler_vector(es, e, es.out_min, k, D_v, sv, G, age, true, m);
                return_euler(c.rc + (src_base + size_t(e.nbr) * src_step) * RC_STRIDE, h, m);
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                y[i] += m[i];
            }
        }
    }
    for (int k = in_s; k < in_e; ++k) {
        in_edge(k);
    }
    if (changed && age == 0) {
        I.fx_changed = 1; // every writer stores 1
    }
    double* yp = c.y + lane_off * NC;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        yp[i] = y[i];
    }
    if (I.c_out_day >= 0) {
        write_phase_end_day<G>(c, I, lane_off, v, age, y);
    }
}

// After a return pass: settle the instance, schedule another pass, or hand it to the sequential replay.
__global__ void fixup_check(LaunchDev ld, InstDev* inst)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= ld.num_inst) {
        return;
    }
    InstDev& I = inst[ld.inst_ids[i]];
    if (!I.c_active || I.c_action != 2 || I.fx_state >= 2) {
        return;
    }
    if (!I.fx_changed) {
        I.fx_state = 3;
        return;
    }
    if (I.fx_sequential || I.fx_pass + 1 >= FX_MAX_PASSES) {
        I.fx_state = 2;
        return;
    }
    I.fx_state   = 1;
    I.fx_pass    = I.fx_pass + 1;
    I.fx_changed = 0;
    I.fx_passes += 1;
    if (ld.cyc) {
        ld.cyc[CYC_PASS0 + I.fx_pass] = 1; // fx_pass < FX_MAX_PASSES here
    }
}

// ------------------------------------------------------------------------------- exact replay
// Sequential return pass of one instance in the reference's edge order with the negative-value fix-up
// (metapopulation_mobility_instant.h:568-603). Fallback for the rare return phases whose fix-up records
// did not settle within FX_MAX_PASSES passes or overflowed the record pool; lanes = age groups.
template <int G>
__device__ __forceinline__ void replay_body(const LaunchDev& ld, InstDev* __restrict__ inst)
{
    const int ii = ld.inst_ids[blockIdx.x];
    InstDev& I   = inst[ii];
    if (!I.c_active || I.c_action != 2 || I.fx_state != 2) {
        return;
    }
    const SegDev& c = ld.segs[I.seg];
    const int a = int(threadIdx.x);
    // restore the pre-return state of every node
    for (int v = 0; v < I.num_nodes; ++v) {
        if (a < G) {
            const size_t off = (seg_node(I, v) * G + a) * NC;
            for (int i = 0; i < NC; ++i) {
                c.y[off + i] = c.ypre[off + i];
            }
        }
    }
    __syncthreads();
    const EdgeStructDev& es = c.estructs[I.estruct];
    const int32_t* seq      = es.seq;
    const EdgeEnt* oute     = es.out_ent;
    const double h          = I.c_step;
    if (a < G) {
        for (int k = 0; k < es.num_edges; ++k) {
            const EdgeEnt e   = oute[seq[3 * k + 2]];
            const size_t soff = seg_node(I, seq[3 * k]) * G + a;
            const size_t doff = seg_node(I, seq[3 * k + 1]) * G + a;
            double row[SNAP_STRIDE];
            load_row(c.snap, soff, row);
            double m[NC];
            traveler_vector(es, e, es.out_min, seq[3 * k + 2], c.out_d ? c.out_d[soff] : 0.0, row, G, a, true, m);
            return_euler(c.rc + doff * RC_STRIDE, h, m);
            double* yd = c.y + doff * NC;
            double* ys = c.y + soff * NC;
            double yl[NC];
            for (int i = 0; i < NC; ++i) {
                yl[i] = yd[i];
            }
            fixup_negative(yl, m);
            for (int i = 0; i < NC; ++i) {
                ys[i] += m[i];
            }
            for (int i = 0; i < NC; ++i) {
                yd[i] -= m[i];
            }
        }
    }
    __syncthreads();
    if (a == 0) {
        I.replays += 1;
        I.fx_state = 3;
    }
    if (I.c_out_day >= 0 && a < G) {
        for (int v = 0; v < I.num_nodes; ++v) {
            const size_t lane = seg_node(I, v) * G + a;
            double yv[NC];
            for (int i = 0; i < NC; ++i) {
                yv[i] = c.y[lane * NC + i];
            }
            write_phase_end_day<G>(c, I, lane, v, a, yv);
        }
    }
}

// ------------------------------------------------------------------------------- kernels
// One kernel per (G, integrator) for launches over segments of one G; the *_mixed kernels serve launch
// tables whose segments have different G (LaunchDev::G == 0): each block reads its segment's G and runs
// the body specialised for it (blocks never span segments).
__device__ __forceinline__ int block_groups(const LaunchDev& ld)
{
    const int lb = xcd_block();
    if (lb >= ld.num_blocks) {
        return 0;
    }
    const int seg = ld.block_seg[lb];
    return seg < 0 ? 0 : ld.segs[seg].G;
}

#define MGO_DISPATCH_G(G_EXPR, CALL)                                                                           \
    switch (G_EXPR) {                                                                                          \
    case 1: { constexpr int GG = 1; CALL; } break;                                                             \
    case 2: { constexpr int GG = 2; CALL; } break;                                                             \
    case 3: { constexpr int GG = 3; CALL; } break;                                                             \
    case 4: { constexpr int GG = 4; CALL; } break;                                                             \
    case 5: { constexpr int GG = 5; CALL; } break;                                                             \
    case 6: { constexpr int GG = 6; CALL; } break;                                                             \
    case 7: { constexpr int GG = 7; CALL; } break;                                                             \
    case 8: { constexpr int GG = 8; CALL; } break;                                                             \
    default: break;                                                                                            \
    }

template <int G, int K, bool NPI>
__global__ void __launch_bounds__(BLOCK) node_phase(LaunchDev ld, InstDev* __restrict__ inst)
{
    __shared__ double wx[BLOCK * (1 + (5 + G - 1) / G)]; // exchange slots, then the stage-season planes
    node_body<G, K, NPI>(ld, inst, wx);
}

template <int K>
__global__ void __launch_bounds__(BLOCK) node_phase_mixed(LaunchDev ld, InstDev* __restrict__ inst)
{
    __shared__ double wx[BLOCK * 6]; // exchange slots and up to five stage-season planes (G = 1)
    MGO_DISPATCH_G(block_groups(ld), (node_body<GG, K>(ld, inst, wx)))
}

template <int G>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) depart_phase(LaunchDev ld,
                                                                                       InstDev* __restrict__ inst)
{
    depart_body<G>(ld, inst);
}

__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) depart_phase_mixed(LaunchDev ld,
                                                                                             InstDev* __restrict__ inst)
{
    MGO_DISPATCH_G(block_groups(ld), (depart_body<GG>(ld, inst)))
}

template <int G, bool FIRST>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) return_phase(LaunchDev ld,
                                                                                       InstDev* __restrict__ inst,
                                                                                       int pass)
{
    return_body<G, FIRST>(ld, inst, pass);
}

template <bool FIRST>
__global__ void __launch_bounds__(BLOCK) return_phase_mixed(LaunchDev ld, InstDev* __restrict__ inst, int pass)
{
    MGO_DISPATCH_G(block_groups(ld), (return_body<GG, FIRST>(ld, inst, pass)))
}

template <int G>
__global__ void __launch_bounds__(WAVE) replay_phase(LaunchDev ld, InstDev* __restrict__ inst)
{
    replay_body<G>(ld, inst);
}

__global__ void __launch_bounds__(WAVE) replay_phase_mixed(LaunchDev ld, InstDev* __restrict__ inst)
{
    const InstDev& I = inst[ld.inst_ids[blockIdx.x]];
    if (!I.c_active || I.c_action != 2 || I.fx_state != 2) {
        return;
    }
    MGO_DISPATCH_G(ld.segs[I.seg].G, (replay_body<GG>(ld, inst)))
}

// ------------------------------------------------------------------------------- launchers

template <int G>
static hipError_t launch_node_g(const LaunchDev& c, InstDev* inst, hipStream_t s)
{
    const int nb = c.num_blocks;
#define MGO_NODE_LAUNCH(KIND)                                                                                  \
    if (c.npi) {                                                                                               \
        hipLaunchKernelGGL((node_phase<G, KIND, true>), dim3(nb), dim3(BLOCK), 0, s, c, inst);                 \
    }                                                                                                          \
    else {                                                                                                     \
        hipLaunchKernelGGL((node_phase<G, KIND, false>), dim3(nb), dim3(BLOCK), 0, s, c, inst);                \
    }
    switch (c.kind) {
    case MGO_INTEGRATOR_CASH_KARP54:
        MGO_NODE_LAUNCH(MGO_INTEGRATOR_CASH_KARP54)
        break;
    case MGO_INTEGRATOR_RKF45:
        MGO_NODE_LAUNCH(MGO_INTEGRATOR_RKF45)
        break;
    case MGO_INTEGRATOR_CASH_KARP54_FIXED:
        MGO_NODE_LAUNCH(MGO_INTEGRATOR_CASH_KARP54_FIXED)
        break;
    default:
        MGO_NODE_LAUNCH(MGO_INTEGRATOR_EULER)
        break;
    }
#undef MGO_NODE_LAUNCH
    return hipGetLastError();
}

template <int G>
static hipError_t launch_edge_g(const LaunchDev& c, InstDev* inst, hipStream_t s)
{
    const int nb = c.num_blocks;
    hipLaunchKernelGGL((depart_phase<G>), dim3(nb), dim3(BLOCK), 0, s, c, inst);
    hipLaunchKernelGGL((return_phase<G, true>), dim3(nb), dim3(BLOCK), 0, s, c, inst, 0);
    hipError_t err = hipGetLastError();
    const int ni   = (c.num_inst + 255) / 256;
    for (int p = 1; p < FX_MAX_PASSES && err == hipSuccess; ++p) {
        hipLaunchKernelGGL(fixup_check, dim3(ni), dim3(256), 0, s, c, inst);
        hipLaunchKernelGGL((return_phase<G, false>), dim3(nb), dim3(BLOCK), 0, s, c, inst, p);
        err = hipGetLastError();
    }
    if (err != hipSuccess) {
        return err;
    }
    hipLaunchKernelGGL(fixup_check, dim3(ni), dim3(256), 0, s, c, inst);
    if (c.num_inst > 0) {
        hipLaunchKernelGGL((replay_phase<G>), dim3(c.num_inst), dim3(WAVE), 0, s, c, inst);
    }
    return hipGetLastError();
}

hipError_t launch_clock(InstDev* inst, int n, int32_t* cyc, hipStream_t s)
{
    hipLaunchKernelGGL(clock_kernel, dim3(1), dim3(1024), 0, s, inst, n, cyc);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- instance slots
// Serving keeps every resident request's instance record in one table of slots. Admission and retirement
// write whole records (table[ids[k]] = src[k]), class rebuilds renumber the segments (table[ids[k]].seg = k);
// one launch each per loop iteration instead of one copy per request.
static_assert(sizeof(InstDev) % 8 == 0, "instance records are copied as 8-byte words");

__global__ void scatter_slots_kernel(InstDev* __restrict__ table, const InstDev* __restrict__ src,
                                     const int32_t* __restrict__ ids, int n)
{
    constexpr int W = int(sizeof(InstDev) / 8);
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(n) * W) {
        return;
    }
    const int k = int(i / W), w = int(i % W);
    reinterpret_cast<uint64_t*>(table + ids[k])[w] = reinterpret_cast<const uint64_t*>(src + k)[w];
}

__global__ void set_segments_kernel(InstDev* __restrict__ table, const int32_t* __restrict__ ids, int n)
{
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < n) {
        table[ids[k]].seg = k;
    }
}

hipError_t launch_scatter_slots(InstDev* table, const InstDev* src, const int32_t* ids, int n, hipStream_t s)
{
    if (n <= 0) {
        return hipSuccess;
    }
    const int64_t words = int64_t(n) * int64_t(sizeof(InstDev) / 8);
    hipLaunchKernelGGL(scatter_slots_kernel, dim3(unsigned((words + 255) / 256)), dim3(256), 0, s, table, src, ids, n);
    return hipGetLastError();
}

hipError_t launch_set_segments(InstDev* table, const int32_t* ids, int n, hipStream_t s)
{
    if (n <= 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL(set_segments_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, table, ids, n);
    return hipGetLastError();
}

static hipError_t launch_node_mixed(const LaunchDev& c, InstDev* inst, hipStream_t s)
{
    const int nb = c.num_blocks;
    switch (c.kind) {
    case MGO_INTEGRATOR_CASH_KARP54:
        hipLaunchKernelGGL((node_phase_mixed<MGO_INTEGRATOR_CASH_KARP54>), dim3(nb), dim3(BLOCK), 0, s, c, inst);
        break;
    case MGO_INTEGRATOR_RKF45:
        hipLaunchKernelGGL((node_phase_mixed<MGO_INTEGRATOR_RKF45>), dim3(nb), dim3(BLOCK), 0, s, c, inst);
        break;
    case MGO_INTEGRATOR_CASH_KARP54_FIXED:
        hipLaunchKernelGGL((node_phase_mixed<MGO_INTEGRATOR_CASH_KARP54_FIXED>), dim3(nb), dim3(BLOCK), 0, s, c, inst);
        break;
    default:
        hipLaunchKernelGGL((node_phase_mixed<MGO_INTEGRATOR_EULER>), dim3(nb), dim3(BLOCK), 0, s, c, inst);
        break;
    }
    return hipGetLastError();
}

static hipError_t launch_edge_mixed(const LaunchDev& c, InstDev* inst, hipStream_t s)
{
    const int nb = c.num_blocks;
    hipLaunchKernelGGL(depart_phase_mixed, dim3(nb), dim3(BLOCK), 0, s, c, inst);
    hipLaunchKernelGGL((return_phase_mixed<true>), dim3(nb), dim3(BLOCK), 0, s, c, inst, 0);
    hipError_t err = hipGetLastError();
    const int ni   = (c.num_inst + 255) / 256;
    for (int p = 1; p < FX_MAX_PASSES && err == hipSuccess; ++p) {
        hipLaunchKernelGGL(fixup_check, dim3(ni), dim3(256), 0, s, c, inst);
        hipLaunchKernelGGL((return_phase_mixed<false>), dim3(nb), dim3(BLOCK), 0, s, c, inst, p);
        err = hipGetLastError();
    }
    if (err != hipSuccess) {
        return err;
    }
    hipLaunchKernelGGL(fixup_check, dim3(ni), dim3(256), 0, s, c, inst);
    if (c.num_inst > 0) {
        hipLaunchKernelGGL(replay_phase_mixed, dim3(c.num_inst), dim3(WAVE), 0, s, c, inst);
    }
    return hipGetLastError();
}

hipError_t launch_node(const LaunchDev& c, InstDev* inst, hipStream_t s)
{
    switch (c.G) {
    case 0:
        return launch_node_mixed(c, inst, s);
    case 1:
        return launch_node_g<1>(c, inst, s);
    case 2:
        return launch_node_g<2>(c, inst, s);
    case 3:
        return launch_node_g<3>(c, inst, s);
    case 4:
        return launch_node_g<4>(c, inst, s);
    case 5:
        return launch_node_g<5>(c, inst, s);
    case 6:
        return launch_node_g<6>(c, inst, s);
    case 7:
        return launch_node_g<7>(c, inst, s);
    case 8:
        return launch_node_g<8>(c, inst, s);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_edge(const LaunchDev& c, InstDev* inst, hipStream_t s)
{
    switch (c.G) {
    case 0:
        return launch_edge_mixed(c, inst, s);
    case 1:
        return launch_edge_g<1>(c, inst, s);
    case 2:
        return launch_edge_g<2>(c, inst, s);
    case 3:
        return launch_edge_g<3>(c, inst, s);
    case 4:
        return launch_edge_g<4>(c, inst, s);
    case 5:
        return launch_edge_g<5>(c, inst, s);
    case 6:
        return launch_edge_g<6>(c, inst, s);
    case 7:
        return launch_edge_g<7>(c, inst, s);
    case 8:
        return launch_edge_g<8>(c, inst, s);
    default:
        return hipErrorInvalidValue;
    }
}

} // namespace mgo

namespace mgo
{
// ------------------------------------------------------------------------------- ensemble analysis
// Post-processing of the daily outputs of several runs of the same shape (cpp/memilio/data/
// analyze_result.h:267-343). A "position" is one (day, node, compartment) value of a run's output; run r's
// output starts at base[r] (doubles). Every kernel reproduces the reference's arithmetic order.

// ensemble_mean (analyze_result.h:289-315): mean = sum over runs, in run order, of value / num_runs
__global__ void ensemble_mean_kernel(const double* __restrict__ out, const int64_t* __restrict__ base, int runs,
                                     int64_t positions, double* __restrict__ mean)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= positions) {
        return;
    }
    const double n = double(runs);
    double acc     = 0.0;
    for (int r = 0; r < runs; ++r) {
        acc += out[base[r] + p] / n;
    }
    mean[p] = acc;
}

// sum_nodes (analyze_result.h:267-287): per run and day, the node values summed in node order
__global__ void sum_nodes_kernel(const double* __restrict__ out, const int64_t* __restrict__ base, int runs,
                                 int days, int nodes, int width, double* __restrict__ sums)
{
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; // (run, day, element)
    if (q >= int64_t(runs) * days * width) {
        return;
    }
    const int e   = int(q % width);
    const int64_t rd = q / width;
    const int d   = int(rd % days);
    const int r   = int(rd / days);
    const double* src = out + base[r] + (int64_t(d) * nodes) * width + e;
    double acc        = 0.0;
    for (int v = 0; v < nodes; ++v) {
        acc += src[int64_t(v) * width];
    }
    sums[q] = acc;
}

// ensemble_percentile (analyze_result.h:316-343): the value at index floor(num_runs * p) of the sorted
// run values of each position. A block sorts the values of `ppb` consecutive positions in LDS (bitonic
// sort over a power-of-two length padded with +inf) and writes the requested order statistics.
__global__ void __launch_bounds__(256) ensemble_select_kernel(const double* __restrict__ out,
                                                               const int64_t* __restrict__ base, int runs, int npad,
                                                               int ppb, int64_t positions,
                                                               const int32_t* __restrict__ ranks, int num_ranks,
                                                               double* __restrict__ sel)
{
    extern __shared__ double lds[];
    const int64_t p0 = int64_t(blockIdx.x) * ppb;
    const int total  = ppb * npad;
    for (int i = int(threadIdx.x); i < total; i += int(blockDim.x)) {
        const int j      = i % ppb; // position within the block (fastest: adjacent addresses)
        const int r      = i / ppb;
        const int64_t p  = p0 + j;
        double v         = INFINITY;
        if (r < runs && p < positions) {
            v = out[base[r] + p];
        }
        lds[j * npad + r] = v;
    }
    __syncthreads();
    const int half = total / 2;
    for (int k = 2; k <= npad; k <<= 1) {
        for (int s = k >> 1; s > 0; s >>= 1) {
            for (int i = int(threadIdx.x); i < half; i += int(blockDim.x)) {
                const int j   = i / (npad / 2);            // segment
                const int t   = i % (npad / 2);            // pair index within the segment
                const int lo  = 2 * s * (t / s) + (t % s); // first element of the pair
                const int hi  = lo + s;
                const bool up = ((lo & k) == 0);
                double* seg   = lds + j * npad;
                const double a = seg[lo], b = seg[hi];
                if ((a > b) == up) {
                    seg[lo] = b;
                    seg[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    for (int i = int(threadIdx.x); i < ppb * num_ranks; i += int(blockDim.x)) {
        const int j     = i % ppb;
        const int q     = i / ppb;
        const int64_t p = p0 + j;
        if (p < positions) {
            sel[int64_t(q) * positions + p] = lds[j * npad + ranks[q]];
        }
    }
}

hipError_t launch_ensemble_mean(const double* out, const int64_t* base, int runs, int64_t positions, double* mean,
                                hipStream_t s)
{
    const int64_t nb = (positions + 255) / 256;
    hipLaunchKernelGGL(ensemble_mean_kernel, dim3(unsigned(nb)), dim3(256), 0, s, out, base, runs, positions, mean);
    return hipGetLastError();
}

hipError_t launch_sum_nodes(const double* out, const int64_t* base, int runs, int days, int nodes, int width,
                            double* sums, hipStream_t s)
{
    const int64_t n  = int64_t(runs) * days * width;
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(sum_nodes_kernel, dim3(unsigned(nb)), dim3(256), 0, s, out, base, runs, days, nodes, width,
                       sums);
    return hipGetLastError();
}

hipError_t launch_ensemble_select(const double* out, const int64_t* base, int runs, int64_t positions,
                                  const int32_t* ranks, int num_ranks, double* sel, hipStream_t s)
{
    int npad = 2;
    while (npad < runs) {
        npad <<= 1;
    }
    const int ppb = std::max(1, 8192 / npad); // 64 KiB of LDS per block
    const size_t lds = size_t(ppb) * npad * sizeof(double);
    const int64_t nb = (positions + ppb - 1) / ppb;
    hipLaunchKernelGGL(ensemble_select_kernel, dim3(unsigned(nb)), dim3(256), lds, s, out, base, runs, npad, ppb,
                       positions, ranks, num_ranks, sel);
    return hipGetLastError();
}

} // namespace mgo
