// oracle.cpp - CPU restatement of MEmilio v2.2.0's Graph-ODE path (test infrastructure: the parity
// checker and the CPU baseline; never linked into the product library).
//
// The reference cannot be built here (its CMake fetches Eigen/Boost/spdlog); this file restates the
// algorithm from the reference sources, keeping the loop structure and the floating-point evaluation
// order. Build with -ffp-contract=off so products and sums round exactly as the reference's default
// x86-64 build does. Citations are relative to the memilio source tree.
//
// Boost.odeint 1.84.0 (controlled_runge_kutta, default_error_checker, default_step_adjuster,
// runge_kutta_cash_karp54) is not part of the reference tree; its published algorithm is restated in
// CashKarpOdeint below and pinned by the reference fixtures that run it
// (cpp/tests/data/secihurd-compare-cashkarp.csv, pycode/.../ode-secihurd-compare.csv,
// cpp/tests/data/graphsimulation-compare.csv).

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <numbers>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc
{

// cpp/memilio/config.h:72-79  Limits<double>::zero_tolerance()
constexpr double kZeroTol = 1e-12;
constexpr double kPi      = std::numbers::pi_v<double>;

using Vec = std::vector<double>;
// DerivFunction (cpp/memilio/math/integrator.h:37-39): dydt = f(y, t)
using DerivFunction = std::function<void(const double* y, double t, double* dydt)>;

// cpp/memilio/math/smoother.h:46-60
inline double smoother_cosine(double x, double xleft, double xright, double yleft, double yright)
{
    if (x <= xleft) {
        return yleft;
    }
    if (x >= xright) {
        return yright;
    }
    return 0.5 * (yleft - yright) * std::cos(kPi / (xright - xleft) * (x - xleft)) + 0.5 * (yleft + yright);
}

// cpp/memilio/math/floating_point.h:39-62
inline bool floating_point_equal(double v1, double v2, double abs_tol, double rel_tol)
{
    double diff = std::abs(v1 - v2);
    return diff <= abs_tol || diff <= std::max(std::abs(v1), std::abs(v2)) * rel_tol;
}

// ------------------------------------------------------------------------------------------ dampings
struct Damping {
    Vec coeffs; // rows*cols, row major
    int level;
    int type;
    double time;
};

// cpp/memilio/epidemiology/damping.h:267-669 (Dampings): sorted list + accumulated cache.
class Dampings
{
public:
    Dampings(int rows, int cols)
        : m_rows(rows)
        , m_cols(cols)
    {
        update_cache();
    }

    // damping.h:602-613 add_ with insert_sorted_replace on key (time, type, level) (utils/stl_util.h:82-96)
    void add(const Damping& d)
    {
        auto less = [](const Damping& a, const Damping& b) {
            if (a.time != b.time) {
                return a.time < b.time;
            }
            if (a.type != b.type) {
                return a.type < b.type;
            }
            return a.level < b.level;
        };
        auto lb = std::lower_bound(m_dampings.begin(), m_dampings.end(), d, less);
        if (lb != m_dampings.end() && !less(d, *lb)) {
            *lb = d;
        }
        else {
            m_dampings.insert(lb, d);
        }
        m_cache_mats.clear();
        m_cache_times.clear();
        update_cache();
    }

    // damping.h:381-396 get_matrix_at(t)(i, j): upper_bound over the cache, cosine ramp over [t_k - 1, t_k]
    double at(double t, int i, int j) const
    {
        auto ub  = std::upper_bound(m_cache_times.begin(), m_cache_times.end(), t);
        size_t k = size_t(ub - m_cache_times.begin());
        return smoother_cosine(t, m_cache_times[k] - 1.0, m_cache_times[k], m_cache_mats[k - 1][i * m_cols + j],
                               m_cache_mats[k][i * m_cols + j]);
    }

    bool empty() const
    {
        return m_dampings.empty();
    }

    // damping.h:336-341 remove(i)
    void remove(size_t i)
    {
        m_dampings.erase(m_dampings.begin() + long(i));
        m_cache_mats.clear();
        m_cache_times.clear();
        update_cache();
    }

    const std::vector<Damping>& list() const
    {
        return m_dampings;
    }
    int rows() const
    {
        return m_rows;
    }
    int cols() const
    {
        return m_cols;
    }

private:
    // damping.h:570-600 update_cache
    void update_cache()
    {
        const size_t n = size_t(m_rows) * m_cols;
        m_cache_mats.push_back(Vec(n, 0.0));
        m_cache_times.push_back(std::numeric_limits<double>::lowest());
        // active_by_type: (index of the damping whose matrix is active, level, type); sum_by_level: (matrix, level)
        std::vector<std::tuple<Vec, int, int>> active_by_type;
        std::vector<std::pair<Vec, int>> sum_by_level;
        for (auto& damping : m_dampings) {
            update_active_dampings(damping, active_by_type, sum_by_level);
            Vec combined = inclusive_exclusive_sum(sum_by_level);
            if (floating_point_equal(damping.time, m_cache_times.back(), 1e-15, 1e-15)) {
                m_cache_mats.back() = combined;
            }
            else {
                m_cache_mats.push_back(combined);
                m_cache_times.push_back(damping.time);
            }
        }
        m_cache_mats.push_back(m_cache_mats.back());
        m_cache_times.push_back(std::numeric_limits<double>::max());
    }

    // damping.h:616-658 update_active_dampings
    static void update_active_dampings(const Damping& damping, std::vector<std::tuple<Vec, int, int>>& active_by_type,
                                       std::vector<std::pair<Vec, int>>& sum_by_level)
    {
        auto it_active = std::find_if(active_by_type.begin(), active_by_type.end(), [&](auto& a) {
            return std::get<1>(a) == damping.level && std::get<2>(a) == damping.type;
        });
        if (it_active != active_by_type.end()) {
            auto& sum = *std::find_if(sum_by_level.begin(), sum_by_level.end(), [&](auto& s) {
                return s.second == damping.level;
            });
            Vec& active = std::get<0>(*it_active);
            for (size_t k = 0; k < sum.first.size(); ++k) {
                sum.first[k] += damping.coeffs[k] - active[k];
            }
            active = damping.coeffs;
        }
        else {
            active_by_type.emplace_back(damping.coeffs, damping.level, damping.type);
            auto it_sum = std::find_if(sum_by_level.begin(), sum_by_level.end(), [&](auto& s) {
                return s.second == damping.level;
            });
            if (it_sum != sum_by_level.end()) {
                for (size_t k = 0; k < it_sum->first.size(); ++k) {
                    it_sum->first[k] += damping.coeffs[k];
                }
            }
            else {
                sum_by_level.emplace_back(damping.coeffs, damping.level);
            }
        }
    }

    // damping.h:549-566 inclusive_exclusive_sum: sum = sum + b - sum*b over levels
    static Vec inclusive_exclusive_sum(const std::vector<std::pair<Vec, int>>& v)
    {
        Vec sum = v.front().first;
        for (size_t l = 1; l < v.size(); ++l) {
            const Vec& b = v[l].first;
            for (size_t k = 0; k < sum.size(); ++k) {
                double prod = sum[k] * b[k];
                sum[k]      = sum[k] + b[k] - prod;
            }
        }
        return sum;
    }

    int m_rows, m_cols;
    std::vector<Damping> m_dampings;
    std::vector<Vec> m_cache_mats;
    Vec m_cache_times;
};

// cpp/memilio/epidemiology/contact_matrix.h:190-194 (DampingMatrixExpression::get_matrix_at)
struct ContactMatrix {
    int G;
    Vec baseline, minimum;
    Dampings dampings;
    ContactMatrix(int g)
        : G(g)
        , baseline(size_t(g) * g, 0.0)
        , minimum(size_t(g) * g, 0.0)
        , dampings(g, g)
    {
    }
    double at(double t, int i, int j) const
    {
        const size_t k = size_t(i) * G + j;
        return baseline[k] - (dampings.at(t, i, j) * (baseline[k] - minimum[k]));
    }
};

// contact_matrix.h:392-400 (ContactMatrixGroup::get_matrix_at): sum over the group's matrices
struct ContactGroup {
    std::vector<ContactMatrix> mats;
    double at(double t, int i, int j) const
    {
        double s = 0.0;
        for (auto& m : mats) {
            s = s + m.at(t, i, j);
        }
        return s;
    }
};

// ------------------------------------------------------------------------------------------ dynamic NPIs
// DampingSampling (epidemiology/damping_sampling.h:40-230) as far as implement_dynamic_npis reads it.
struct NpiDamping {
    double value;
    int level, type;
    std::vector<int> mats; // matrix indices of the group
    Vec weights;           // group weights
};

// make_contact_damping_matrix (damping_sampling.h:282-288): d_ij = 1 - sqrt((1 - g_i)(1 - g_j)), g = value * w
static Vec make_contact_damping_matrix(const Vec& g)
{
    const size_t n = g.size();
    Vec m(n * n);
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < n; ++j) {
            m[i * n + j] = 1.0 - std::sqrt((1.0 - g[i]) * (1.0 - g[j]));
        }
    }
    return m;
}

// make_mobility_damping_vector (damping_sampling.h:298-310): [g_0 (rows / G times), g_1, ...]
static Vec make_mobility_damping_vector(size_t rows, const Vec& g)
{
    Vec v(rows);
    const size_t per = rows / g.size();
    for (size_t k = 0; k < rows; ++k) {
        v[k] = g[k / per];
    }
    return v;
}

// get_active_damping (dynamic_npis.h:291-303): coefficients of the last damping of (level, type) with
// time <= t, zeros if none
static Vec get_active_damping(const Dampings& d, int level, int type, double t)
{
    const auto& l = d.list();
    for (auto it = l.rbegin(); it != l.rend(); ++it) {
        if (it->level == level && it->type == type && it->time <= t) {
            return it->coeffs;
        }
    }
    return Vec(size_t(d.rows()) * d.cols(), 0.0);
}

// get_damping_indices (dynamic_npis.h:260-272): dampings of (level, type) with begin < time < end
static std::vector<size_t> get_damping_indices(const Dampings& d, int level, int type, double begin, double end)
{
    std::vector<size_t> idx;
    const auto& l = d.list();
    for (size_t i = 0; i < l.size(); ++i) {
        if (l[i].level == level && l[i].type == type && l[i].time > begin && l[i].time < end) {
            idx.push_back(i);
        }
    }
    return idx;
}

// implement_dynamic_npis (dynamic_npis.h:318-392) on the dampings of a group of matrices. make_matrix:
// 0 = the scaled weight vector itself (reshaped), 1 = make_contact_damping_matrix.
static void implement_dynamic_npis(std::vector<Dampings*>& group, const std::vector<NpiDamping>& npis, double begin,
                                   double end, int make_matrix)
{
    for (auto& npi : npis) {
        for (int mat_idx : npi.mats) {
            Dampings& expr = *group[size_t(mat_idx)];
            Vec active     = get_active_damping(expr, npi.level, npi.type, begin);
            Vec active_end = get_active_damping(expr, npi.level, npi.type, end);
            Vec g(npi.weights.size());
            for (size_t k = 0; k < g.size(); ++k) {
                g[k] = npi.value * npi.weights[k];
            }
            Vec value = make_matrix == 1 ? make_contact_damping_matrix(g)
                        : make_matrix == 2 ? make_mobility_damping_vector(size_t(expr.rows()), g)
                                           : g;
            bool implemented = false;
            bool any_less    = false;
            for (size_t k = 0; k < value.size(); ++k) {
                any_less = any_less || (active[k] < value[k]);
            }
            if (any_less) {
                Damping d{Vec(value.size()), npi.level, npi.type, begin};
                for (size_t k = 0; k < value.size(); ++k) {
                    d.coeffs[k] = std::max(value[k], active[k]);
                }
                expr.add(d);
                implemented = true;
            }
            for (size_t i : get_damping_indices(expr, npi.level, npi.type, begin, end)) {
                const Damping& cur = expr.list()[i];
                Damping d{Vec(value.size()), npi.level, npi.type, cur.time};
                for (size_t k = 0; k < value.size(); ++k) {
                    d.coeffs[k] = std::max(cur.coeffs[k], value[k]);
                }
                expr.add(d);
                implemented = true;
            }
            if (implemented) {
                expr.add(Damping{active_end, npi.level, npi.type, end});
            }
        }
    }
    // remove a damping whose previous damping of the same (level, type) has the same coefficients
    for (Dampings* expr : group) {
        for (int i = 0; i < int(expr->list().size()) - 1; ++i) {
            const auto& l    = expr->list();
            const size_t cur = l.size() - 1 - size_t(i);
            for (size_t p = cur; p-- > 0;) {
                if (l[p].level == l[cur].level && l[p].type == l[cur].type) {
                    if (l[p].coeffs == l[cur].coeffs) {
                        expr->remove(cur);
                    }
                    break;
                }
            }
        }
    }
}

// DynamicNPIs (epidemiology/dynamic_npis.h:36-242)
struct DynamicNPIs {
    std::vector<std::pair<double, std::vector<NpiDamping>>> thresholds; // sorted by value, descending
    double duration = 14.0, delay = 0.0, base = 1.0;
    double begin = std::numeric_limits<double>::lowest(), end = std::numeric_limits<double>::max();
    // set_threshold: insert_sorted_replace with t1.first > t2.first (utils/stl_util.h:82-96)
    void set_threshold(double v, const std::vector<NpiDamping>& d)
    {
        auto it = thresholds.begin();
        while (it != thresholds.end() && it->first > v) {
            ++it;
        }
        if (it != thresholds.end() && !(v > it->first)) {
            it->second = d;
        }
        else {
            thresholds.insert(it, std::make_pair(v, d));
        }
    }
    // get_max_exceeded_threshold: first threshold (descending) below value; -1 if none
    int max_exceeded(double value) const
    {
        for (size_t k = 0; k < thresholds.size(); ++k) {
            if (value > thresholds[k].first) {
                return int(k);
            }
        }
        return -1;
    }
};

// floating_point_greater_equal(v1, v2, abs_tol) (math/floating_point.h:96-100), rel_tol = DBL_MIN
inline bool floating_point_greater_equal(double v1, double v2, double abs_tol)
{
    const double diff = v1 - v2;
    return !(diff < -(abs_tol + std::numeric_limits<double>::min() * std::max(std::abs(v1), std::abs(v2))));
}

// ------------------------------------------------------------------------------------------ models
struct Model {
    virtual ~Model() = default;
    virtual int num_groups() const                                 = 0;
    virtual int num_comparts() const                               = 0;
    virtual std::unique_ptr<Model> clone() const                   = 0;
    // FlowModel::get_derivatives(pop, y, t, dydt) (cpp/memilio/compartments/flow_model.h:138-144)
    virtual void get_derivatives(const double* pop, const double* y, double t, double* dydt) const = 0;
    // get_mobility_factors customisation point (metapopulation_mobility_instant.h:507-519): default ones
    virtual void mobility_factors(const double* /*y*/, double* factors) const
    {
        std::fill(factors, factors + size(), 1.0);
    }
    // test_commuters customisation point (metapopulation_mobility_instant.h:531-542): default no-op
    virtual void test_commuters(double* /*from_last_value*/, double* /*mobile*/, double /*t*/) const
    {
    }
    int size() const
    {
        return num_groups() * num_comparts();
    }
    Vec initial;
};

enum Comp { cS = 0, cE, cINS, cINSC, cISy, cISyC, cISev, cICr, cR, cD, NC };

// cpp/models/ode_secir/model.h:45-59 (Flows order), :82-214 (get_flows)
struct OsecirModel : Model {
    int G;
    double start_day = 0, seasonality = 0, icu_capacity = std::numeric_limits<double>::max(),
           ttc = std::numeric_limits<double>::max(), ttc_max_risk = 5.0;
    double nondetection = 0, start_detection = 0, end_detection = 0;
    Vec T_E, T_INS, T_ISy, T_ISev, T_ICr, tp, rel, R, risk, max_risk, sev, cps, dps, dpc;
    ContactGroup contacts;
    DynamicNPIs npis; // DynamicNPIsInfectedSymptoms (models/ode_secir/parameters.h:342-355)
    mutable Vec flows;

    explicit OsecirModel(int g)
        : G(g)
        , flows(size_t(15) * g, 0.0)
    {
    }
    int num_groups() const override
    {
        return G;
    }
    int num_comparts() const override
    {
        return NC;
    }
    std::unique_ptr<Model> clone() const override
    {
        return std::make_unique<OsecirModel>(*this);
    }

    void get_flows(const double* pop, const double* y, double t, double* f) const
    {
        double icu_occupancy           = 0.0;
        double test_and_trace_required = 0.0;
        for (int i = 0; i < G; ++i) {
            test_and_trace_required += (1 - R[i]) / T_INS[i] * pop[i * NC + cINS];
            icu_occupancy += pop[i * NC + cICr];
        }
        for (int i = 0; i < G; ++i) {
            const int Si = i * NC + cS, Ei = i * NC + cE, INSi = i * NC + cINS, INSCi = i * NC + cINSC,
                      ISyi = i * NC + cISy, ISyCi = i * NC + cISyC, ISevi = i * NC + cISev, ICri = i * NC + cICr;
            double* fi = f + 15 * i;
            for (int j = 0; j < G; ++j) {
                const int Sj = j * NC + cS, Ej = j * NC + cE, INSj = j * NC + cINS, ISyj = j * NC + cISy,
                          ISevj = j * NC + cISev, ICrj = j * NC + cICr, Rj = j * NC + cR;
                double riskFromInfectedSymptomatic =
                    smoother_cosine(test_and_trace_required, ttc, ttc * ttc_max_risk, risk[j], max_risk[j]);
                double season_val    = (1 + seasonality * std::sin(kPi * ((start_day + t) / 182.5 + 0.5)));
                double cont_freq_eff = season_val * contacts.at(t, i, j);
                double Nj            = pop[Sj] + pop[Ej] + pop[INSj] + pop[ISyj] + pop[ISevj] + pop[ICrj] + pop[Rj];
                const double divNj   = (Nj < kZeroTol) ? 0.0 : 1.0 / Nj;
                double dummy_S       = y[Si] * cont_freq_eff * divNj * tp[i] *
                                 (rel[j] * pop[INSj] + riskFromInfectedSymptomatic * pop[ISyj]);
                fi[0] += dummy_S;
            }
            double criticalPerSevereAdjusted = smoother_cosine(icu_occupancy, 0.90 * icu_capacity, icu_capacity, cps[i], 0.0);
            double deathsPerSevereAdjusted   = cps[i] - criticalPerSevereAdjusted;

            fi[1]  = (1.0 / T_E[i]) * y[Ei];
            fi[2]  = (1 - R[i]) * (1 / T_INS[i]) * y[INSi];
            fi[3]  = R[i] * (1.0 / T_INS[i]) * y[INSi];
            fi[4]  = (1.0 - R[i]) * (1.0 / T_INS[i]) * y[INSCi];
            fi[5]  = R[i] * (1.0 / T_INS[i]) * y[INSCi];
            fi[6]  = sev[i] / T_ISy[i] * y[ISyi];
            fi[7]  = (1 - sev[i]) / T_ISy[i] * y[ISyi];
            fi[8]  = sev[i] / T_ISy[i] * y[ISyCi];
            fi[9]  = (1.0 - sev[i]) / T_ISy[i] * y[ISyCi];
            fi[10] = criticalPerSevereAdjusted / T_ISev[i] * y[ISevi];
            fi[11] = (1.0 - cps[i] - dps[i]) / T_ISev[i] * y[ISevi];
            fi[12] = (dps[i] + deathsPerSevereAdjusted) / T_ISev[i] * y[ISevi];
            fi[13] = dpc[i] / T_ICr[i] * y[ICri];
            fi[14] = (1.0 - dpc[i]) / T_ICr[i] * y[ICri];
        }
    }

    void get_derivatives(const double* pop, const double* y, double t, double* dydt) const override
    {
        std::fill(flows.begin(), flows.end(), 0.0);
        get_flows(pop, y, t, flows.data());
        // flow_model.h:112-125,219-235: dydt[source] -= f; dydt[target] += f, flows in Flows order
        static const int src[15] = {cS, cE, cINS, cINS, cINSC, cINSC, cISy, cISy, cISyC, cISyC, cISev, cISev, cISev, cICr, cICr};
        static const int tgt[15] = {cE, cINS, cISy, cR, cISyC, cR, cISev, cR, cISev, cR, cICr, cR, cD, cD, cR};
        std::fill(dydt, dydt + size(), 0.0);
        for (int i = 0; i < G; ++i) {
            for (int k = 0; k < 15; ++k) {
                dydt[i * NC + src[k]] -= flows[15 * i + k];
                dydt[i * NC + tgt[k]] += flows[15 * i + k];
            }
        }
    }

    // cpp/models/ode_secir/model.h:687-715 get_mobility_factors
    void mobility_factors(const double* y, double* factors) const override
    {
        double test_and_trace_required = 0;
        for (int g = 0; g < G; ++g) {
            double term = (1 - R[g]) / T_INS[g] * y[g * NC + cINS];
            test_and_trace_required = (g == 0) ? term : test_and_trace_required + term;
        }
        std::fill(factors, factors + size(), 1.0);
        for (int g = 0; g < G; ++g) {
            factors[g * NC + cISy] = smoother_cosine(test_and_trace_required, ttc, ttc * ttc_max_risk, risk[g], max_risk[g]);
        }
    }

    // cpp/models/ode_secir/model.h:717-742 test_commuters
    void test_commuters(double* last, double* mobile, double time) const override
    {
        double nondet = 1.0;
        if (time >= start_detection && time < end_detection) {
            nondet = nondetection;
        }
        for (int i = 0; i < G; ++i) {
            const int INSi = i * NC + cINS, INSCi = i * NC + cINSC, ISyi = i * NC + cISy, ISyCi = i * NC + cISyC;
            last[INSi] -= mobile[INSi] * (1 - nondet);
            last[INSCi] += mobile[INSi] * (1 - nondet);
            last[ISyi] -= mobile[ISyi] * (1 - nondet);
            last[ISyCi] += mobile[ISyi] * (1 - nondet);
            mobile[ISyi] *= nondet;
            mobile[INSi] *= nondet;
        }
    }
};

// cpp/models/ode_seir/model.h:47-102 (S, E, I, R; flows S->E, E->I, I->R)
struct OseirModel : Model {
    int G;
    Vec T_E, T_I, tp;
    ContactGroup contacts;
    mutable Vec flows;
    explicit OseirModel(int g)
        : G(g)
        , flows(size_t(3) * g, 0.0)
    {
    }
    int num_groups() const override
    {
        return G;
    }
    int num_comparts() const override
    {
        return 4;
    }
    std::unique_ptr<Model> clone() const override
    {
        return std::make_unique<OseirModel>(*this);
    }
    void get_derivatives(const double* pop, const double* y, double t, double* dydt) const override
    {
        std::fill(flows.begin(), flows.end(), 0.0);
        for (int i = 0; i < G; ++i) {
            const int Si = 4 * i, Ei = 4 * i + 1, Ii = 4 * i + 2;
            for (int j = 0; j < G; ++j) {
                const int Sj = 4 * j, Ej = 4 * j + 1, Ij = 4 * j + 2, Rj = 4 * j + 3;
                const double Nj        = pop[Sj] + pop[Ej] + pop[Ij] + pop[Rj];
                const double divNj     = (Nj < kZeroTol) ? 0.0 : 1.0 / Nj;
                const double coeffStoE = contacts.at(t, i, j) * tp[i] * divNj;
                flows[3 * i] += coeffStoE * y[Si] * pop[Ij];
            }
            flows[3 * i + 1] = (1.0 / T_E[i]) * y[Ei];
            flows[3 * i + 2] = (1.0 / T_I[i]) * y[Ii];
        }
        std::fill(dydt, dydt + size(), 0.0);
        for (int i = 0; i < G; ++i) {
            dydt[4 * i + 0] -= flows[3 * i];
            dydt[4 * i + 1] += flows[3 * i];
            dydt[4 * i + 1] -= flows[3 * i + 1];
            dydt[4 * i + 2] += flows[3 * i + 1];
            dydt[4 * i + 2] -= flows[3 * i + 2];
            dydt[4 * i + 3] += flows[3 * i + 2];
        }
    }
};

// ------------------------------------------------------------------------------------------ integrators
// IntegratorCore (cpp/memilio/math/integrator.h:46-129)
struct Core {
    double dt_min, dt_max;
    int64_t attempts = 0;
    Core(double mn, double mx)
        : dt_min(mn)
        , dt_max(mx)
    {
    }
    virtual ~Core()                                                                                   = default;
    virtual std::unique_ptr<Core> clone() const                                                       = 0;
    virtual bool step(const DerivFunction& f, const Vec& yt, double& t, double& dt, Vec& ytp1) = 0;
};

// Cash-Karp 5(4) coefficients as odeint defines them (runge_kutta_cash_karp54.hpp: each a ratio of two
// integers evaluated in double). Cash & Karp, ACM TOMS 16 (1990) 201-222.
struct CashKarpTableau {
    double c[6], a[6][5], b[6], db[6];
    CashKarpTableau()
    {
        std::memset(this, 0, sizeof(*this));
        c[0] = 0.0;
        c[1] = 1.0 / 5.0;
        c[2] = 3.0 / 10.0;
        c[3] = 3.0 / 5.0;
        c[4] = 1.0;
        c[5] = 7.0 / 8.0;
        a[1][0] = 1.0 / 5.0;
        a[2][0] = 3.0 / 40.0;
        a[2][1] = 9.0 / 40.0;
        a[3][0] = 3.0 / 10.0;
        a[3][1] = -9.0 / 10.0;
        a[3][2] = 6.0 / 5.0;
        a[4][0] = -11.0 / 54.0;
        a[4][1] = 5.0 / 2.0;
        a[4][2] = -70.0 / 27.0;
        a[4][3] = 35.0 / 27.0;
        a[5][0] = 1631.0 / 55296.0;
        a[5][1] = 175.0 / 512.0;
        a[5][2] = 575.0 / 13824.0;
        a[5][3] = 44275.0 / 110592.0;
        a[5][4] = 253.0 / 4096.0;
        b[0] = 37.0 / 378.0;
        b[1] = 0.0;
        b[2] = 250.0 / 621.0;
        b[3] = 125.0 / 594.0;
        b[4] = 0.0;
        b[5] = 512.0 / 1771.0;
        db[0] = 37.0 / 378.0 - 2825.0 / 27648.0;
        db[1] = 0.0;
        db[2] = 250.0 / 621.0 - 18575.0 / 48384.0;
        db[3] = 125.0 / 594.0 - 13525.0 / 55296.0;
        db[4] = -277.0 / 14336.0;
        db[5] = 512.0 / 1771.0 - 1.0 / 4.0;
    }
};
static const CashKarpTableau kCK;

// odeint explicit_error_generic_rk::do_step_impl with the generic_rk_algorithm stage sums:
// x_tmp = 1*x + (a_s0*dt)*k1 + ... ; out = 1*x + (b0*dt)*k1 + ... ; xerr = (db0*dt)*k1 + ...
static void cash_karp_do_step(const DerivFunction& f, const Vec& x, const Vec& dxdt, double t, Vec& out, double dt,
                              Vec* xerr, std::vector<Vec>& F, Vec& x_tmp)
{
    const size_t n = x.size();
    // F[0] = dxdt (k1); F[1..5] = k2..k6
    for (int s = 1; s < 6; ++s) {
        for (size_t i = 0; i < n; ++i) {
            double v = x[i] + (kCK.a[s][0] * dt) * dxdt[i];
            for (int m = 1; m < s; ++m) {
                v = v + (kCK.a[s][m] * dt) * F[m][i];
            }
            x_tmp[i] = v;
        }
        f(x_tmp.data(), t + kCK.c[s] * dt, F[s].data());
    }
    for (size_t i = 0; i < n; ++i) {
        double v = x[i] + (kCK.b[0] * dt) * dxdt[i];
        for (int m = 1; m < 6; ++m) {
            v = v + (kCK.b[m] * dt) * F[m][i];
        }
        out[i] = v;
    }
    if (xerr) {
        for (size_t i = 0; i < n; ++i) {
            double v = (kCK.db[0] * dt) * dxdt[i];
            for (int m = 1; m < 6; ++m) {
                v = v + (kCK.db[m] * dt) * F[m][i];
            }
            (*xerr)[i] = v;
        }
    }
}

// ControlledStepperWrapper<FP, runge_kutta_cash_karp54> (cpp/memilio/math/stepper_wrapper.h:56-202) with
// odeint's controlled_runge_kutta::try_step, default_error_checker (a_x = a_dxdt = 1) and
// default_step_adjuster (max_dt = dt_max).
struct CashKarpOdeint : Core {
    double abs_tol, rel_tol;
    Vec dxdt, xerr, x_tmp;
    std::vector<Vec> F;
    CashKarpOdeint(double a, double r, double mn, double mx)
        : Core(mn, mx)
        , abs_tol(a)
        , rel_tol(r)
    {
    }
    std::unique_ptr<Core> clone() const override
    {
        return std::make_unique<CashKarpOdeint>(abs_tol, rel_tol, dt_min, dt_max);
    }

    // controlled_runge_kutta::try_step(system, in, t, out, dt)
    bool try_step(const DerivFunction& f, const Vec& in, double& t, Vec& out, double& dt)
    {
        const size_t n = in.size();
        if (dxdt.size() != n) {
            dxdt.assign(n, 0.0);
            xerr.assign(n, 0.0);
            x_tmp.assign(n, 0.0);
            F.assign(6, Vec(n, 0.0));
        }
        f(in.data(), t, dxdt.data());
        ++attempts;
        // default_step_adjuster::check_step_size_limit: less_eq_with_sign(dt, max_dt, dt)
        if (dt_max != 0.0 && !(dt - dt_max <= std::numeric_limits<double>::epsilon())) {
            dt = dt_max;
            return false;
        }
        cash_karp_do_step(f, in, dxdt, t, out, dt, &xerr, F, x_tmp);
        // default_error_checker::error: max_i |xerr_i| / (eps_abs + eps_rel*(a_x*|x_i| + a_dxdt*dt*|dxdt_i|))
        const double a_dxdt_dt = 1.0 * dt;
        double max_rel_err     = 0.0;
        for (size_t i = 0; i < n; ++i) {
            double e = std::abs(xerr[i]) / (abs_tol + rel_tol * (1.0 * std::abs(in[i]) + a_dxdt_dt * std::abs(dxdt[i])));
            max_rel_err = std::max(max_rel_err, e);
        }
        if (max_rel_err > 1.0) {
            // default_step_adjuster::decrease_step, error_order 4
            dt *= std::max(9.0 / 10.0 * std::pow(max_rel_err, -1.0 / (4 - 1)), 1.0 / 5.0);
            if (dt_max != 0.0) {
                dt = std::min(dt, dt_max);
            }
            return false;
        }
        t += dt;
        // default_step_adjuster::increase_step, stepper_order 5
        if (max_rel_err < 0.5) {
            max_rel_err = std::max(std::pow(5.0, -5.0), max_rel_err);
            dt *= 9.0 / 10.0 * std::pow(max_rel_err, -1.0 / 5);
            if (dt_max != 0.0) {
                dt = std::min(dt, dt_max);
            }
        }
        return true;
    }

    // stepper_wrapper.h:112-163
    bool step(const DerivFunction& f, const Vec& yt, double& t, double& dt, Vec& ytp1) override
    {
        bool fail        = true;
        bool is_dt_valid = true;
        while (fail && is_dt_valid) {
            if (dt < dt_min) {
                is_dt_valid = false;
                dt          = dt_min;
            }
            fail = !try_step(f, yt, t, ytp1, dt);
        }
        dt = std::max(dt, dt_min);
        if (fail) {
            t += dt_min;
            return false;
        }
        return true;
    }
};

// ExplicitStepperWrapper<FP, runge_kutta_cash_karp54> (stepper_wrapper.h:208-259)
struct CashKarpFixed : Core {
    Vec dxdt, x_tmp;
    std::vector<Vec> F;
    CashKarpFixed()
        : Core(0.0, 0.0)
    {
    }
    std::unique_ptr<Core> clone() const override
    {
        return std::make_unique<CashKarpFixed>();
    }
    bool step(const DerivFunction& f, const Vec& yt, double& t, double& dt, Vec& ytp1) override
    {
        const size_t n = yt.size();
        if (dxdt.size() != n) {
            dxdt.assign(n, 0.0);
            x_tmp.assign(n, 0.0);
            F.assign(6, Vec(n, 0.0));
        }
        f(yt.data(), t, dxdt.data());
        ++attempts;
        cash_karp_do_step(f, yt, dxdt, t, ytp1, dt, nullptr, F, x_tmp);
        t += dt;
        return true;
    }
};

// RKIntegratorCore with Tableau / TableauFinal (cpp/memilio/math/adapt_rk.h:62-310), default RKF45
struct RKF45 : Core {
    double abs_tol, rel_tol;
    std::vector<Vec> tab; // tab[i] = (c, a_1 .. a_{i+1})
    Vec b_low, b_high;
    std::vector<Vec> K; // K[col][i]
    Vec yt_eval;
    RKF45(double a, double r, double mn, double mx)
        : Core(mn, mx)
        , abs_tol(a)
        , rel_tol(r)
    {
        tab.resize(5);
        for (size_t i = 0; i < tab.size(); ++i) {
            tab[i].resize(i + 2);
        }
        tab[0][0] = 0.25;
        tab[0][1] = 0.25;
        tab[1][0] = 3 / 8.0;
        tab[1][1] = 3 / 32.0;
        tab[1][2] = 9 / 32.0;
        tab[2][0] = 12 / 13.0;
        tab[2][1] = 1932 / 2197.0;
        tab[2][2] = -7200 / 2197.0;
        tab[2][3] = 7296 / 2197.0;
        tab[3][0] = 1.0;
        tab[3][1] = 439 / 216.0;
        tab[3][2] = -8.0;
        tab[3][3] = 3680 / 513.0;
        tab[3][4] = -845 / 4104.0;
        tab[4][0] = 0.5;
        tab[4][1] = -8 / 27.0;
        tab[4][2] = 2.0;
        tab[4][3] = -3544 / 2565.0;
        tab[4][4] = 1859 / 4104.0;
        tab[4][5] = -11 / 40.0;
        b_low  = {25 / 216.0, 0.0, 1408 / 2565.0, 2197 / 4104.0, -0.2, 0.0};
        b_high = {16 / 135.0, 0.0, 6656 / 12825.0, 28561 / 56430.0, -9 / 50.0, 2 / 55.0};
    }
    std::unique_ptr<Core> clone() const override
    {
        return std::make_unique<RKF45>(abs_tol, rel_tol, dt_min, dt_max);
    }

    // adapt_rk.h:217-298
    bool step(const DerivFunction& f, const Vec& yt, double& t, double& dt, Vec& ytp1) override
    {
        const size_t n  = yt.size();
        const size_t nk = b_low.size();
        dt              = std::min(dt, dt_max);
        if (yt_eval.size() != n) {
            yt_eval.assign(n, 0.0);
            K.assign(nk, Vec(n, 0.0));
        }
        yt_eval = yt;
        bool converged = false, dt_is_invalid = false;
        Vec err(n), eps(n), bdiff(nk);
        for (size_t k = 0; k < nk; ++k) {
            bdiff[k] = b_high[k] - b_low[k];
        }
        while (!converged && !dt_is_invalid) {
            if (dt < dt_min) {
                dt_is_invalid = true;
                dt            = dt_min;
            }
            ++attempts;
            f(yt_eval.data(), t, K[0].data());
            for (size_t i = 1; i < nk; ++i) {
                double t_eval = t;
                t_eval += tab[i - 1][0] * dt;
                ytp1 = yt_eval;
                for (size_t k = 1; k < tab[i - 1].size(); ++k) {
                    const double c = dt * tab[i - 1][k];
                    for (size_t j = 0; j < n; ++j) {
                        ytp1[j] += c * K[k - 1][j];
                    }
                }
                f(ytp1.data(), t_eval, K[i].data());
            }
            // ytp1 = y + dt*(K*b_low): column-major GEMV sums the columns in order, then scales
            for (size_t j = 0; j < n; ++j) {
                double s = 0.0;
                for (size_t k = 0; k < nk; ++k) {
                    s = s + K[k][j] * b_low[k];
                }
                ytp1[j] = yt_eval[j] + s * dt;
            }
            for (size_t j = 0; j < n; ++j) {
                double s = 0.0;
                for (size_t k = 0; k < nk; ++k) {
                    s = s + K[k][j] * bdiff[k];
                }
                err[j] = dt * std::abs(s);
                eps[j] = abs_tol + std::abs(ytp1[j]) * rel_tol;
            }
            converged = true;
            for (size_t j = 0; j < n; ++j) {
                converged = converged && (err[j] <= eps[j]);
            }
            if (converged || dt_is_invalid) {
                t += dt;
            }
            double mn = std::numeric_limits<double>::infinity();
            for (size_t j = 0; j < n; ++j) {
                mn = std::min(mn, eps[j] / err[j]);
            }
            double dt_new = dt * std::pow(mn, (1. / (nk - 1)));
            dt_new *= 0.9;
            dt = std::min(dt_new, dt_max);
        }
        dt = std::max(dt, dt_min);
        return converged;
    }
};

// EulerIntegratorCore (cpp/memilio/math/euler.h:32-63)
struct Euler : Core {
    Euler()
        : Core(0.0, 0.0)
    {
    }
    std::unique_ptr<Core> clone() const override
    {
        return std::make_unique<Euler>();
    }
    bool step(const DerivFunction& f, const Vec& yt, double& t, double& dt, Vec& ytp1) override
    {
        f(yt.data(), t, ytp1.data());
        ++attempts;
        for (size_t i = 0; i < yt.size(); ++i) {
            ytp1[i] = yt[i] + dt * ytp1[i];
        }
        t += dt;
        return true;
    }
};

static std::unique_ptr<Core> make_core(const mgo_integrator_cfg& c)
{
    switch (c.kind) {
    case MGO_INTEGRATOR_CASH_KARP54:
        return std::make_unique<CashKarpOdeint>(c.abs_tol, c.rel_tol, c.dt_min, c.dt_max);
    case MGO_INTEGRATOR_RKF45:
        return std::make_unique<RKF45>(c.abs_tol, c.rel_tol, c.dt_min, c.dt_max);
    case MGO_INTEGRATOR_CASH_KARP54_FIXED:
        return std::make_unique<CashKarpFixed>();
    case MGO_INTEGRATOR_EULER:
        return std::make_unique<Euler>();
    default:
        return nullptr;
    }
}

// ------------------------------------------------------------------------------------------ time series
// cpp/memilio/utils/time_series.h:47-914 (only what the path uses)
struct TimeSeries {
    int n;
    Vec times;
    Vec values;
    explicit TimeSeries(int n_)
        : n(n_)
    {
    }
    int size() const
    {
        return int(times.size());
    }
    void add_time_point(double t, const double* v)
    {
        times.push_back(t);
        values.insert(values.end(), v, v + n);
    }
    void add_time_point(double t)
    {
        times.push_back(t);
        values.resize(values.size() + n, 0.0);
    }
    double* value(int i)
    {
        return values.data() + size_t(i) * n;
    }
    const double* value(int i) const
    {
        return values.data() + size_t(i) * n;
    }
    double* last_value()
    {
        return value(size() - 1);
    }
    double& last_time()
    {
        return times.back();
    }
    void remove_time_point(int i)
    {
        times.erase(times.begin() + i);
        values.erase(values.begin() + size_t(i) * n, values.begin() + size_t(i + 1) * n);
    }
    // time_series.h:905-914 find_value_reverse: index of the last point with time == t_search (tolerances)
    int find_reverse(double t_search, double abs_tol, double rel_tol) const
    {
        for (int i = size() - 1; i >= 0; --i) {
            if (floating_point_equal(times[i], t_search, abs_tol, rel_tol)) {
                return i;
            }
        }
        return -1;
    }
};

// ------------------------------------------------------------------------------------------ simulation
// SystemIntegrator::advance (cpp/memilio/math/integrator.h:184-249)
static void system_integrator_advance(Core& core, const DerivFunction& f, double tmax, double& dt, TimeSeries& results,
                                      int64_t* steps)
{
    const double t0  = results.times.back();
    const double eps = kZeroTol;
    double dt_restore     = 0.0;
    double dt_min_restore = core.dt_min;
    double t              = t0;
    Vec ycur(results.n), ynext(results.n);
    for (int i = results.size() - 1; std::abs((tmax - t) / (tmax - t0)) > eps; ++i) {
        if (dt > tmax - t) {
            dt_restore  = dt;
            dt          = tmax - t;
            core.dt_min = std::min(tmax - t, core.dt_min);
        }
        std::copy(results.value(i), results.value(i) + results.n, ycur.begin());
        core.step(f, ycur, t, dt, ynext);
        results.add_time_point(t, ynext.data());
        if (steps) {
            ++*steps;
        }
    }
    core.dt_min = dt_min_restore;
    dt          = std::max(dt, dt_restore);
}

// mio::Simulation / SimulationBase (cpp/memilio/compartments/simulation.h:41-74, simulation_base.h:40-176)
struct Simulation {
    std::unique_ptr<Model> model;
    std::unique_ptr<Core> core;
    TimeSeries result;
    double dt;
    int64_t steps = 0;
    Simulation(const Model& m, std::unique_ptr<Core> c, double t0, double dt_)
        : model(m.clone())
        , core(std::move(c))
        , result(m.size())
        , dt(dt_)
    {
        result.add_time_point(t0, model->initial.data());
    }
    bool frozen = false;                 // the reference tests' MockSimulation: state constant over time
    double npi_value = std::numeric_limits<double>::lowest(), npi_end = 0.0; // m_dynamic_npi
    // mio::Simulation::advance: RHS = model.eval_right_hand_side(y, y, t, dydt)
    void base_advance(double tmax)
    {
        if (frozen) {
            result.add_time_point(tmax, result.last_value());
            return;
        }
        const Model* m = model.get();
        DerivFunction f = [m](const double* y, double t, double* dydt) {
            m->get_derivatives(y, y, t, dydt);
        };
        system_integrator_advance(*core, f, tmax, dt, result, &steps);
    }
    // osecir::Simulation::advance (models/ode_secir/model.h:289-367): dynamic NPIs checked at the start of
    // every piece of at most one day, pieces clamped at the end of the active NPI
    void advance(double tmax)
    {
        auto* om = dynamic_cast<OsecirModel*>(model.get());
        if (!om || om->npis.thresholds.empty()) {
            base_advance(tmax);
            return;
        }
        DynamicNPIs& dn = om->npis;
        double t        = result.times.back();
        while (t < tmax) {
            const double delay = t > 0 ? dn.delay : 0.0;
            if (floating_point_greater_equal(t, dn.begin, 1e-10) && t < dn.end) {
                const double inf_rel = infections_relative(*om, result.last_value()) * dn.base;
                const int k          = dn.max_exceeded(inf_rel);
                const bool expired   = floating_point_greater_equal(t, npi_end, 1e-10);
                if (k >= 0 && (dn.thresholds[size_t(k)].first > npi_value || expired)) {
                    const bool renewal = expired && !(dn.thresholds[size_t(k)].first > npi_value);
                    const double eff   = renewal ? 0.0 : delay;
                    if (t + eff < dn.end) {
                        const double t_start = t + eff;
                        const double t_end   = std::min(dn.end, t_start + dn.duration);
                        npi_value            = dn.thresholds[size_t(k)].first;
                        npi_end              = t_end;
                        const double tsd     = (!renewal && t_start > 0.0) ? t_start + 1.0 : t_start;
                        std::vector<Dampings*> group;
                        for (auto& cm : om->contacts.mats) {
                            group.push_back(&cm.dampings);
                        }
                        implement_dynamic_npis(group, dn.thresholds[size_t(k)].second, tsd, t_end + 1.0, 1);
                    }
                }
            }
            double dt_eff = std::min(1.0, tmax - t);
            if (t < npi_end) {
                dt_eff = std::min(dt_eff, npi_end - t);
            }
            base_advance(t + dt_eff);
            t = t + dt_eff;
        }
    }
    // get_infections_relative (models/ode_secir/model.h:413-423): sum of InfectedSymptoms over the age
    // groups / the model's total population
    static double infections_relative(const OsecirModel& m, const double* y)
    {
        double sum_inf = 0.0;
        for (int g = 0; g < m.G; ++g) {
            sum_inf += y[g * NC + cISy];
        }
        double total = 0.0;
        for (double v : m.initial) {
            total += v;
        }
        return sum_inf / total;
    }
};

// SimulationNode (cpp/memilio/mobility/metapopulation_mobility_instant.h:39-101)
struct SimulationNode {
    Simulation sim;
    Vec last_state;
    SimulationNode(const Model& m, std::unique_ptr<Core> c, double t0, double dt)
        : sim(m, std::move(c), t0, dt)
        , last_state(sim.result.last_value(), sim.result.last_value() + sim.result.n)
    {
    }
    void advance(double t, double dt)
    {
        sim.advance(t + dt);
        std::copy(sim.result.last_value(), sim.result.last_value() + sim.result.n, last_state.begin());
    }
};

// MobilityEdge with constant coefficients (metapopulation_mobility_instant.h:295-623); edge dynamic NPIs
// and saved commuter indices are not on this path.
struct MobilityEdge {
    Vec coeffs;         // baseline B of the MobilityCoefficients
    Vec minimum;        // minimum M
    Dampings dampings;  // vector-shaped dampings of the coefficients
    DynamicNPIs npis;   // MobilityParameters::get_dynamic_npis_infected
    double npi_value = std::numeric_limits<double>::lowest(), npi_end = 0.0; // m_dynamic_npi
    TimeSeries mobile;
    Vec return_times;
    bool return_mobile_population = false;
    explicit MobilityEdge(const Vec& c)
        : coeffs(c)
        , minimum(c.size(), 0.0)
        , dampings(int(c.size()), 1)
        , mobile(int(c.size()))
    {
    }

    // DampingMatrixExpression::get_matrix_at (contact_matrix.h:190-194) of the coefficient vector
    Vec coeffs_at(double t) const
    {
        Vec c(coeffs.size());
        for (size_t j = 0; j < c.size(); ++j) {
            c[j] = coeffs[j] - (dampings.at(t, int(j), 0) * (coeffs[j] - minimum[j]));
        }
        return c;
    }

    void apply_mobility(double t, double dt, SimulationNode& from, SimulationNode& to, orc_stats* st)
    {
        const int n = mobile.n;
        // edge dynamic NPIs (:549-566): checked on the source's last state at every call
        if (!npis.thresholds.empty()) {
            auto* om = dynamic_cast<const OsecirModel*>(from.sim.model.get());
            const double inf_rel = Simulation::infections_relative(*om, from.last_state.data()) * npis.base;
            const int k          = npis.max_exceeded(inf_rel);
            const bool expired   = floating_point_greater_equal(t, npi_end, 1e-10);
            if (k >= 0 && (npis.thresholds[size_t(k)].first > npi_value || expired)) {
                const double t_end = t + npis.duration;
                npi_value          = npis.thresholds[size_t(k)].first;
                npi_end            = t_end;
                std::vector<Dampings*> group{&dampings};
                implement_dynamic_npis(group, npis.thresholds[size_t(k)].second, t, t_end, 2);
            }
        }
        // returns (:568-603)
        for (int i = mobile.size() - 1; i >= 0; --i) {
            if (return_times[size_t(i)] <= t) {
                const int iv = to.sim.result.find_reverse(mobile.times[size_t(i)], 1e-10, 1e-10);
                const double* v0 = to.sim.result.value(iv);
                double* m        = mobile.value(i);
                // calculate_mobility_returns (:460-471): one Euler step of the travelers with the
                // destination model, pop = destination value at the departure time
                Vec y0(m, m + n), dydt(n);
                to.sim.model->get_derivatives(v0, y0.data(), mobile.times[size_t(i)], dydt.data());
                for (int j = 0; j < n; ++j) {
                    m[j] = y0[j] + dt * dydt[j];
                }
                if (st) {
                    st->rhs_evals++;
                    st->returns++;
                }
                // negative fix-up (:576-596)
                double* to_last = to.sim.result.last_value();
                Vec remaining(n);
                for (int j = 0; j < n; ++j) {
                    remaining[j] = to_last[j] - m[j];
                }
                const int nc = to.sim.model->num_comparts();
                for (int j = 0; j < n; ++j) {
                    if (remaining[j] < 0) {
                        const int group = j / nc;
                        int max_index   = 0;
                        for (int k = 1; k < nc; ++k) {
                            if (remaining[group * nc + k] > remaining[group * nc + max_index]) {
                                max_index = k;
                            }
                        }
                        max_index += group * nc;
                        m[max_index] -= remaining[j];
                        m[j] += remaining[j];
#ifdef ORC_TRACE_FIXUPS
                        fprintf(stderr, "fixup t=%g comp=%d group=%d rem=%.3e to_last=%.3e\n", t, j % nc, group,
                                remaining[j], to_last[j]);
#endif
                        if (st) {
                            st->negative_fixups++;
                        }
                    }
                }
                double* from_last = from.sim.result.last_value();
                for (int j = 0; j < n; ++j) {
                    from_last[j] += m[j];
                }
                for (int j = 0; j < n; ++j) {
                    to_last[j] -= m[j];
                }
                mobile.remove_time_point(i);
                return_times.erase(return_times.begin() + i);
            }
        }
        // departures (:605-621)
        const Vec cat     = coeffs_at(t);
        bool any_positive = false;
        for (double c : cat) {
            any_positive = any_positive || (c > 0.0);
        }
        if (!return_mobile_population && any_positive) {
            Vec factors(n);
            from.sim.model->mobility_factors(from.last_state.data(), factors.data());
            Vec m(n);
            for (int j = 0; j < n; ++j) {
                m[j] = from.last_state[j] * cat[j] * factors[j];
            }
            mobile.add_time_point(t, m.data());
            return_times.push_back(t + dt);
            from.sim.model->test_commuters(from.sim.result.last_value(), mobile.last_value(), t);
            double* to_last   = to.sim.result.last_value();
            double* from_last = from.sim.result.last_value();
            const double* ml  = mobile.last_value();
            for (int j = 0; j < n; ++j) {
                to_last[j] += ml[j];
            }
            for (int j = 0; j < n; ++j) {
                from_last[j] -= ml[j];
            }
        }
        return_mobile_population = !return_mobile_population;
    }
};

struct EdgeEntry {
    int start, end;
    MobilityEdge edge;
};

// GraphSimulation with make_mobility_sim node/edge functions (graph_simulation.h:91-121,
// metapopulation_mobility_instant.h:629-678)
struct GraphSim {
    double t, dt;
    std::vector<SimulationNode> nodes;
    std::vector<EdgeEntry> edges; // sorted by (start, end) as Graph::add_edge keeps them (graph.h:231-241)
    orc_stats* st = nullptr;
    int node_threads = 1; // > 1: the node phase of each half-day runs on that many OpenMP threads

    void add_edge(int s, int e, const Vec& c)
    {
        add_edge(s, e, MobilityEdge(c));
    }
    void add_edge(int s, int e, const MobilityEdge& me)
    {
        EdgeEntry item{s, e, me};
        auto less = [](const EdgeEntry& a, const EdgeEntry& b) {
            return a.start == b.start ? a.end < b.end : a.start < b.start;
        };
        auto range = std::equal_range(edges.begin(), edges.end(), item, less);
        if (range.second - range.first == 1) {
            *range.first = item;
        }
        else {
            edges.insert(range.first, item);
        }
    }

    void advance(double t_max)
    {
        double step = dt;
        while (t < t_max) {
            if (t + step > t_max) {
                step = t_max - t;
            }
            if (node_threads > 1) {
                // nodes advance independently (each owns its model, integrator and TimeSeries); the edges
                // below stay sequential in the reference's order
#pragma omp parallel for schedule(dynamic, 4) num_threads(node_threads)
                for (int k = 0; k < int(nodes.size()); ++k) {
                    nodes[size_t(k)].advance(t, step);
                }
            }
            else {
                for (auto& n : nodes) {
                    n.advance(t, step);
                }
            }
            t += step;
            for (auto& e : edges) {
                e.edge.apply_mobility(t, step, nodes[size_t(e.start)], nodes[size_t(e.end)], st);
            }
        }
    }
};

// interpolate_simulation_result (cpp/memilio/data/analyze_result.h:186-265), abs_tol = 100*zero_tolerance
static void interpolate_daily(const TimeSeries& ts, std::vector<double>& days_out, std::vector<Vec>& vals_out)
{
    const double abs_tol = 100.0 * kZeroTol;
    const double t0      = ts.times.front();
    const double t_max   = ts.times.back();
    const int day0       = int(std::ceil(t0 - abs_tol));
    const int day_max    = int(std::floor(t_max + abs_tol));
    std::vector<double> tps;
    for (int d = day0; d <= day_max; ++d) {
        tps.push_back(double(d));
    }
    days_out.clear();
    vals_out.clear();
    const int n     = ts.n;
    size_t interp_idx = 0;
    if (tps.empty()) {
        return;
    }
    if (ts.times[0] >= tps[0]) {
        days_out.push_back(tps[0]);
        vals_out.emplace_back(ts.value(0), ts.value(0) + n);
        ++interp_idx;
    }
    for (int sim_idx = 0; sim_idx < ts.size() - 1 && interp_idx < tps.size();) {
        const double ta = ts.times[size_t(sim_idx)], tb = ts.times[size_t(sim_idx) + 1];
        if (ta < tps[interp_idx] && tb >= tps[interp_idx]) {
            // math/interpolation.h:44-48 linear_interpolation
            const double weight = (tps[interp_idx] - ta) / (tb - ta);
            const double* ya    = ts.value(sim_idx);
            const double* yb    = ts.value(sim_idx + 1);
            Vec v(n);
            for (int k = 0; k < n; ++k) {
                v[k] = ya[k] + weight * (yb[k] - ya[k]);
            }
            days_out.push_back(tps[interp_idx]);
            vals_out.push_back(v);
            ++interp_idx;
        }
        else {
            ++sim_idx;
        }
    }
    if (interp_idx < tps.size() && ts.times.back() < tps[interp_idx]) {
        days_out.push_back(tps[interp_idx]);
        vals_out.emplace_back(ts.value(ts.size() - 1), ts.value(ts.size() - 1) + n);
    }
}

// ------------------------------------------------------------------------------------------ descriptors
static bool build_osecir(const mgo_graph_desc* g, int k, OsecirModel& m)
{
    const int G = g->num_groups;
    const double* sc = g->node_scalars + size_t(k) * MGO_NUM_SCALARS;
    m.start_day       = sc[MGO_P_START_DAY];
    m.seasonality     = sc[MGO_P_SEASONALITY];
    m.icu_capacity    = sc[MGO_P_ICU_CAPACITY];
    m.ttc             = sc[MGO_P_TEST_AND_TRACE_CAPACITY];
    m.ttc_max_risk    = sc[MGO_P_TEST_AND_TRACE_MAX_RISK];
    m.nondetection    = sc[MGO_P_COMMUTER_NONDETECTION];
    m.start_detection = sc[MGO_P_START_COMMUTER_DETECTION];
    m.end_detection   = sc[MGO_P_END_COMMUTER_DETECTION];
    auto age = [&](int a) {
        const double* p = g->age_params + (size_t(k) * MGO_NUM_AGE_PARAMS + a) * G;
        return Vec(p, p + G);
    };
    m.T_E      = age(MGO_A_TIME_EXPOSED);
    m.T_INS    = age(MGO_A_TIME_INFECTED_NO_SYMPTOMS);
    m.T_ISy    = age(MGO_A_TIME_INFECTED_SYMPTOMS);
    m.T_ISev   = age(MGO_A_TIME_INFECTED_SEVERE);
    m.T_ICr    = age(MGO_A_TIME_INFECTED_CRITICAL);
    m.tp       = age(MGO_A_TRANSMISSION_PROBABILITY_ON_CONTACT);
    m.rel      = age(MGO_A_RELATIVE_TRANSMISSION_NO_SYMPTOMS);
    m.R        = age(MGO_A_RECOVERED_PER_INFECTED_NO_SYMPTOMS);
    m.risk     = age(MGO_A_RISK_OF_INFECTION_FROM_SYMPTOMATIC);
    m.max_risk = age(MGO_A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC);
    m.sev      = age(MGO_A_SEVERE_PER_INFECTED_SYMPTOMS);
    m.cps      = age(MGO_A_CRITICAL_PER_SEVERE);
    m.dps      = age(MGO_A_DEATHS_PER_SEVERE);
    m.dpc      = age(MGO_A_DEATHS_PER_CRITICAL);
    ContactMatrix cm(G);
    const size_t gg = size_t(G) * G;
    std::copy(g->contact_baseline + k * gg, g->contact_baseline + (k + 1) * gg, cm.baseline.begin());
    if (g->contact_minimum) {
        std::copy(g->contact_minimum + k * gg, g->contact_minimum + (k + 1) * gg, cm.minimum.begin());
    }
    if (g->damping_offsets) {
        for (int d = g->damping_offsets[k]; d < g->damping_offsets[k + 1]; ++d) {
            Damping dm;
            dm.coeffs.assign(g->damping_coeffs + size_t(d) * gg, g->damping_coeffs + size_t(d + 1) * gg);
            dm.level = g->damping_levels ? g->damping_levels[d] : 0;
            dm.type  = g->damping_types ? g->damping_types[d] : 0;
            dm.time  = g->damping_times[d];
            cm.dampings.add(dm);
        }
    }
    m.contacts.mats.clear();
    m.contacts.mats.push_back(cm);
    if (const mgo_dynamic_npis* dn = g->dynamic_npis) {
        const double* par = dn->node_params + size_t(k) * MGO_NPI_NUM_PARAMS;
        m.npis.duration   = par[MGO_NPI_DURATION];
        m.npis.delay      = par[MGO_NPI_DELAY];
        m.npis.base       = par[MGO_NPI_BASE_VALUE];
        m.npis.begin      = par[MGO_NPI_DIRECTIVE_BEGIN];
        m.npis.end        = par[MGO_NPI_DIRECTIVE_END];
        for (int t = dn->node_offsets[k]; t < dn->node_offsets[k + 1]; ++t) {
            std::vector<NpiDamping> ds;
            for (int d = dn->damping_offsets[t]; d < dn->damping_offsets[t + 1]; ++d) {
                NpiDamping nd;
                nd.value = dn->damping_values[d];
                nd.level = dn->damping_levels[d];
                nd.type  = dn->damping_types[d];
                nd.mats  = {0};
                nd.weights.assign(dn->damping_weights + size_t(d) * G, dn->damping_weights + size_t(d + 1) * G);
                ds.push_back(nd);
            }
            m.npis.set_threshold(dn->thresholds[t], ds);
        }
    }
    const double* p = g->populations + size_t(k) * G * NC;
    m.initial.assign(p, p + size_t(G) * NC);
    return true;
}

static bool valid_desc(const mgo_graph_desc* g)
{
    return g && g->num_nodes >= 1 && g->num_groups >= 1 && g->node_scalars && g->age_params &&
           g->contact_baseline && g->populations && (g->num_edges == 0 || (g->edge_start && g->edge_end && g->edge_coeffs));
}

static void build_graph(const mgo_graph_desc* g, const mgo_sim_cfg* cfg, GraphSim& sim)
{
    sim.t  = cfg->t0;
    sim.dt = cfg->dt_graph;
    sim.nodes.reserve(size_t(g->num_nodes));
    for (int k = 0; k < g->num_nodes; ++k) {
        OsecirModel m(g->num_groups);
        build_osecir(g, k, m);
        sim.nodes.emplace_back(m, make_core(cfg->integrator), cfg->t0, cfg->dt_node);
    }
    const size_t n   = size_t(g->num_groups) * NC;
    const size_t len = (g->edge_coeff_len == 1) ? 1 : n;
    const int G = g->num_groups;
    for (int e = 0; e < g->num_edges; ++e) {
        // MobilityEdge(Eigen::VectorX) coefficients; a single value stands for VectorX::Constant(n, c)
        Vec c = (len == 1) ? Vec(n, g->edge_coeffs[e])
                           : Vec(g->edge_coeffs + size_t(e) * n, g->edge_coeffs + size_t(e + 1) * n);
        MobilityEdge me(c);
        if (g->edge_coeff_minimum) {
            me.minimum = (len == 1) ? Vec(n, g->edge_coeff_minimum[e])
                                    : Vec(g->edge_coeff_minimum + size_t(e) * n,
                                          g->edge_coeff_minimum + size_t(e + 1) * n);
        }
        // the source node's out-edge dampings (per age group, make_mobility_damping_vector) and NPIs
        const int src = g->edge_start[e];
        if (g->out_damping_offsets) {
            for (int k = g->out_damping_offsets[src]; k < g->out_damping_offsets[src + 1]; ++k) {
                Vec gv(g->out_damping_coeffs + size_t(k) * G, g->out_damping_coeffs + size_t(k + 1) * G);
                me.dampings.add(Damping{make_mobility_damping_vector(n, gv), g->out_damping_levels[k],
                                        g->out_damping_types[k], g->out_damping_times[k]});
            }
        }
        if (const mgo_dynamic_npis* dn = g->out_dynamic_npis) {
            const double* par = dn->node_params + size_t(src) * MGO_NPI_NUM_PARAMS;
            me.npis.duration  = par[MGO_NPI_DURATION];
            me.npis.base      = par[MGO_NPI_BASE_VALUE];
            for (int t = dn->node_offsets[src]; t < dn->node_offsets[src + 1]; ++t) {
                std::vector<NpiDamping> ds;
                for (int d = dn->damping_offsets[t]; d < dn->damping_offsets[t + 1]; ++d) {
                    NpiDamping nd;
                    nd.value = dn->damping_values[d];
                    nd.level = dn->damping_levels[d];
                    nd.type  = dn->damping_types[d];
                    nd.mats  = {0};
                    nd.weights.assign(dn->damping_weights + size_t(d) * G, dn->damping_weights + size_t(d + 1) * G);
                    ds.push_back(nd);
                }
                me.npis.set_threshold(dn->thresholds[t], ds);
            }
        }
        sim.add_edge(g->edge_start[e], g->edge_end[e], me);
    }
}

static int32_t run_graph(const mgo_graph_desc* g, const mgo_sim_cfg* cfg, double tmax, int32_t capacity_days,
                         double* daily_out, int32_t* num_days, orc_stats* stats, int node_threads = 1)
{
    if (!valid_desc(g) || !cfg || !make_core(cfg->integrator)) {
        return MGO_ERR_INVALID_VALUE;
    }
    GraphSim sim;
    orc_stats local{};
    sim.st           = &local;
    sim.node_threads = node_threads;
    build_graph(g, cfg, sim);
    sim.advance(tmax);
    const size_t n = size_t(g->num_groups) * NC;
    int32_t nd     = 0;
    for (int k = 0; k < g->num_nodes; ++k) {
        std::vector<double> days;
        std::vector<Vec> vals;
        interpolate_daily(sim.nodes[size_t(k)].sim.result, days, vals);
        nd = int32_t(days.size());
        for (int32_t d = 0; d < nd && d < capacity_days; ++d) {
            if (daily_out) {
                std::copy(vals[size_t(d)].begin(), vals[size_t(d)].end(),
                          daily_out + (size_t(d) * g->num_nodes + k) * n);
            }
        }
        local.steps += sim.nodes[size_t(k)].sim.steps;
        local.attempts += sim.nodes[size_t(k)].sim.core->attempts;
    }
    local.rhs_evals += 6 * local.attempts;
    if (num_days) {
        *num_days = nd;
    }
    if (stats) {
        *stats = local;
    }
    return MGO_OK;
}

} // namespace orc

using namespace orc;

extern "C" {

int32_t orc_simulate(const mgo_graph_desc* model, const mgo_sim_cfg* cfg, double tmax, int32_t capacity, double* times,
                     double* values)
{
    if (!valid_desc(model) || !cfg) {
        return MGO_ERR_INVALID_VALUE;
    }
    auto core = make_core(cfg->integrator);
    if (!core) {
        return -MGO_ERR_INVALID_VALUE;
    }
    OsecirModel m(model->num_groups);
    build_osecir(model, 0, m);
    Simulation sim(m, std::move(core), cfg->t0, cfg->dt_node);
    sim.advance(tmax);
    const int np = sim.result.size();
    for (int i = 0; i < np && i < capacity; ++i) {
        times[i] = sim.result.times[size_t(i)];
        std::copy(sim.result.value(i), sim.result.value(i) + sim.result.n, values + size_t(i) * sim.result.n);
    }
    return np;
}

int32_t orc_simulate_daily(const mgo_graph_desc* model, const mgo_sim_cfg* cfg, double tmax, int32_t capacity_days,
                           double* daily_out, orc_stats* stats)
{
    if (!valid_desc(model) || !cfg) {
        return -MGO_ERR_INVALID_VALUE;
    }
    auto core = make_core(cfg->integrator);
    if (!core) {
        return -MGO_ERR_INVALID_VALUE;
    }
    OsecirModel m(model->num_groups);
    build_osecir(model, 0, m);
    Simulation sim(m, std::move(core), cfg->t0, cfg->dt_node);
    sim.advance(tmax);
    std::vector<double> days;
    std::vector<Vec> vals;
    interpolate_daily(sim.result, days, vals);
    for (size_t d = 0; d < days.size() && int32_t(d) < capacity_days; ++d) {
        std::copy(vals[d].begin(), vals[d].end(), daily_out + d * sim.result.n);
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->steps    = sim.steps;
        stats->attempts = sim.core->attempts;
    }
    return int32_t(days.size());
}

int32_t orc_graph_run(const mgo_graph_desc* graph, const mgo_sim_cfg* cfg, double tmax, int32_t capacity_days,
                      double* daily_out, int32_t* num_days, orc_stats* stats)
{
    return run_graph(graph, cfg, tmax, capacity_days, daily_out, num_days, stats);
}

int32_t orc_graph_run_threads(const mgo_graph_desc* graph, const mgo_sim_cfg* cfg, double tmax, int32_t node_threads,
                              int32_t capacity_days, double* daily_out, int32_t* num_days, orc_stats* stats)
{
    return run_graph(graph, cfg, tmax, capacity_days, daily_out, num_days, stats, node_threads < 1 ? 1 : node_threads);
}

int32_t orc_graph_run_node_series(const mgo_graph_desc* graph, const mgo_sim_cfg* cfg, double tmax, int32_t node,
                                  int32_t capacity, double* times, double* values)
{
    if (!valid_desc(graph) || !cfg || node < 0 || node >= graph->num_nodes || !make_core(cfg->integrator)) {
        return -MGO_ERR_INVALID_VALUE;
    }
    GraphSim sim;
    build_graph(graph, cfg, sim);
    sim.advance(tmax);
    const TimeSeries& ts = sim.nodes[size_t(node)].sim.result;
    for (int i = 0; i < ts.size() && i < capacity; ++i) {
        times[i] = ts.times[size_t(i)];
        std::copy(ts.value(i), ts.value(i) + ts.n, values + size_t(i) * ts.n);
    }
    return ts.size();
}

int32_t orc_ensemble_run(const mgo_ensemble_desc* desc, int32_t threads, double* daily_out, orc_stats* stats)
{
    if (!desc || desc->num_runs < 0) {
        return MGO_ERR_INVALID_VALUE;
    }
    const int nr = desc->num_runs;
    std::vector<size_t> offset(size_t(nr) + 1, 0);
    for (int r = 0; r < nr; ++r) {
        const mgo_graph_desc& g = desc->graphs[r];
        const int ndays = int(std::floor(desc->tmax[r] + 1e-10)) - int(std::ceil(desc->cfgs[r].t0 - 1e-10)) + 1;
        offset[size_t(r) + 1] = offset[size_t(r)] + size_t(ndays) * g.num_nodes * g.num_groups * NC;
    }
    std::vector<orc_stats> st(static_cast<size_t>(nr));
    int32_t status = MGO_OK;
#ifdef _OPENMP
    if (threads < 1) {
        threads = omp_get_max_threads();
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
    for (int r = 0; r < nr; ++r) {
        const mgo_graph_desc& g = desc->graphs[r];
        const int ndays = int((offset[size_t(r) + 1] - offset[size_t(r)]) / (size_t(g.num_nodes) * g.num_groups * NC));
        int32_t nd = 0;
        int32_t s  = run_graph(&g, &desc->cfgs[r], desc->tmax[r], ndays, daily_out ? daily_out + offset[size_t(r)] : nullptr,
                               &nd, &st[size_t(r)]);
        if (s != MGO_OK) {
#ifdef _OPENMP
#pragma omp critical
#endif
            status = s;
        }
    }
    if (stats) {
        orc_stats tot{};
        for (auto& s : st) {
            tot.steps += s.steps;
            tot.attempts += s.attempts;
            tot.rhs_evals += s.rhs_evals;
            tot.negative_fixups += s.negative_fixups;
            tot.returns += s.returns;
        }
        *stats = tot;
    }
    (void)threads;
    return status;
}

int32_t orc_oseir_graph_series(int32_t num_nodes, int32_t G, const double* params, const double* contact,
                               const double* populations, int32_t num_edges, const int32_t* edge_start,
                               const int32_t* edge_end, const double* edge_coeffs, double t0, double dt_graph,
                               double dt_node, double tmax, int32_t node, int32_t capacity, double* times,
                               double* values)
{
    if (num_nodes < 1 || G < 1 || node < 0 || node >= num_nodes) {
        return -MGO_ERR_INVALID_VALUE;
    }
    GraphSim sim;
    sim.t  = t0;
    sim.dt = dt_graph;
    for (int k = 0; k < num_nodes; ++k) {
        OseirModel m(G);
        const double* p = params + size_t(k) * 3 * G;
        m.T_E.assign(p, p + G);
        m.T_I.assign(p + G, p + 2 * G);
        m.tp.assign(p + 2 * G, p + 3 * G);
        ContactMatrix cm(G);
        std::copy(contact + size_t(k) * G * G, contact + size_t(k + 1) * G * G, cm.baseline.begin());
        m.contacts.mats.push_back(cm);
        m.initial.assign(populations + size_t(k) * G * 4, populations + size_t(k + 1) * G * 4);
        // default core of mio::Simulation: ControlledStepperWrapper<cash_karp54>(1e-10, 1e-5, DBL_MIN, DBL_MAX)
        sim.nodes.emplace_back(m,
                               std::make_unique<CashKarpOdeint>(1e-10, 1e-5, std::numeric_limits<double>::min(),
                                                                std::numeric_limits<double>::max()),
                               t0, dt_node);
    }
    for (int e = 0; e < num_edges; ++e) {
        Vec c(edge_coeffs + size_t(e) * 4 * G, edge_coeffs + size_t(e + 1) * 4 * G);
        sim.add_edge(edge_start[e], edge_end[e], c);
    }
    sim.advance(tmax);
    const TimeSeries& ts = sim.nodes[size_t(node)].sim.result;
    for (int i = 0; i < ts.size() && i < capacity; ++i) {
        times[i] = ts.times[size_t(i)];
        std::copy(ts.value(i), ts.value(i) + ts.n, values + size_t(i) * ts.n);
    }
    return ts.size();
}

int32_t orc_edge_scenario(const mgo_graph_desc* model, double t, const double* coeffs, double* out)
{
    if (!valid_desc(model)) {
        return MGO_ERR_INVALID_VALUE;
    }
    OsecirModel m(model->num_groups);
    build_osecir(model, 0, m);
    auto core = [] {
        return std::make_unique<CashKarpOdeint>(1e-10, 1e-5, std::numeric_limits<double>::min(),
                                                std::numeric_limits<double>::max());
    };
    SimulationNode node1(m, core(), t, 0.1), node2(m, core(), t, 0.1);
    const int n = m.size();
    MobilityEdge edge(Vec(coeffs, coeffs + n));
    auto dump = [&](int k) {
        std::copy(node1.sim.result.last_value(), node1.sim.result.last_value() + n, out + size_t(2 * k) * n);
        std::copy(node2.sim.result.last_value(), node2.sim.result.last_value() + n, out + size_t(2 * k + 1) * n);
    };
    edge.apply_mobility(t, 0.5, node1, node2, nullptr);
    dump(0);
    node1.advance(t, 0.5);
    node2.advance(t, 0.5);
    t += 0.5;
    edge.apply_mobility(t, 0.5, node1, node2, nullptr);
    dump(1);
    node1.advance(t, 0.5);
    node2.advance(t, 0.5);
    t += 0.5;
    edge.apply_mobility(t, 0.5, node1, node2, nullptr);
    dump(2);
    return MGO_OK;
}

int32_t orc_integrate_test_problem(int32_t problem, const mgo_integrator_cfg* cfg, double t0, double tmax, double dt,
                                   int32_t capacity, double* times, double* values)
{
    auto core = make_core(*cfg);
    if (!core || problem != 0) {
        return -MGO_ERR_INVALID_VALUE;
    }
    DerivFunction f = [](const double* /*y*/, double t, double* dydt) {
        dydt[0] = std::cos(t);
    };
    TimeSeries ts(1);
    double y0 = std::sin(t0);
    ts.add_time_point(t0, &y0);
    system_integrator_advance(*core, f, tmax, dt, ts, nullptr);
    for (int i = 0; i < ts.size() && i < capacity; ++i) {
        times[i]  = ts.times[size_t(i)];
        values[i] = ts.value(i)[0];
    }
    return ts.size();
}

// implement_dynamic_npis on a group of `num_mats` damping lists of rows x cols matrices (the known-answer
// scenarios of cpp/tests/test_dynamic_npis.cpp:145-317). Lists in/out as CSR over the matrices: entry e has
// times[e], levels[e], types[e], coeffs[e][rows*cols]. npis: value, level, type, matrix indices
// (npi_mat_offsets CSR into npi_mats) and weights[num_weights]. make_matrix: 0 = the scaled weights,
// 1 = make_contact_damping_matrix. Returns the number of output entries (or a negative status).
int32_t orc_implement_npis(int32_t rows, int32_t cols, int32_t num_mats, const int32_t* list_offsets,
                           const double* times, const int32_t* levels, const int32_t* types, const double* coeffs,
                           int32_t num_npis, const double* npi_values, const int32_t* npi_levels,
                           const int32_t* npi_types, const int32_t* npi_mat_offsets, const int32_t* npi_mats,
                           int32_t num_weights, const double* npi_weights, double begin, double end,
                           int32_t make_matrix, int32_t capacity, int32_t* out_offsets, double* out_times,
                           int32_t* out_levels, int32_t* out_types, double* out_coeffs)
{
    const size_t rc = size_t(rows) * cols;
    std::vector<Dampings> mats;
    for (int m = 0; m < num_mats; ++m) {
        mats.emplace_back(rows, cols);
        for (int e = list_offsets[m]; e < list_offsets[m + 1]; ++e) {
            mats.back().add(Damping{Vec(coeffs + e * rc, coeffs + (e + 1) * rc), levels[e], types[e], times[e]});
        }
    }
    std::vector<NpiDamping> npis;
    for (int k = 0; k < num_npis; ++k) {
        NpiDamping d;
        d.value = npi_values[k];
        d.level = npi_levels[k];
        d.type  = npi_types[k];
        d.mats.assign(npi_mats + npi_mat_offsets[k], npi_mats + npi_mat_offsets[k + 1]);
        d.weights.assign(npi_weights + size_t(k) * num_weights, npi_weights + size_t(k + 1) * num_weights);
        npis.push_back(d);
    }
    std::vector<Dampings*> group;
    for (auto& m : mats) {
        group.push_back(&m);
    }
    implement_dynamic_npis(group, npis, begin, end, make_matrix);
    int32_t n = 0;
    out_offsets[0] = 0;
    for (int m = 0; m < num_mats; ++m) {
        for (const Damping& d : mats[size_t(m)].list()) {
            if (n < capacity) {
                out_times[n]  = d.time;
                out_levels[n] = d.level;
                out_types[n]  = d.type;
                std::copy(d.coeffs.begin(), d.coeffs.end(), out_coeffs + size_t(n) * rc);
            }
            ++n;
        }
        out_offsets[m + 1] = n;
    }
    return n;
}

// The edge dynamic-NPI check of MobilityEdge::apply_mobility (metapopulation_mobility_instant.h:549-566) over
// a sequence of calls (times[c], infections relative of the source inf_rel[c]; the known-answer scenario of
// cpp/tests/test_dynamic_npis.cpp:319-376): one threshold of one damping (value, level 0, type 0, weights
// [num_weights]) on a coefficient vector of n entries. Writes the number of dampings after each call.
int32_t orc_edge_npi_calls(int32_t n, double threshold, double base, double duration, double value,
                           int32_t num_weights, const double* weights, int32_t num_calls, const double* times,
                           const double* inf_rel, int32_t* out_num_dampings)
{
    Dampings d(n, 1);
    DynamicNPIs npis;
    npis.duration = duration;
    npis.base     = base;
    NpiDamping s;
    s.value = value;
    s.level = 0;
    s.type  = 0;
    s.mats  = {0};
    s.weights.assign(weights, weights + num_weights);
    npis.set_threshold(threshold, {s});
    double npi_value = std::numeric_limits<double>::lowest(), npi_end = 0.0;
    for (int c = 0; c < num_calls; ++c) {
        const double t  = times[c];
        const int k     = npis.max_exceeded(inf_rel[c] * npis.base);
        const bool exp_ = floating_point_greater_equal(t, npi_end, 1e-10);
        if (k >= 0 && (npis.thresholds[size_t(k)].first > npi_value || exp_)) {
            npi_value = npis.thresholds[size_t(k)].first;
            npi_end   = t + npis.duration;
            std::vector<Dampings*> group{&d};
            implement_dynamic_npis(group, npis.thresholds[size_t(k)].second, t, npi_end, 2);
        }
        out_num_dampings[c] = int32_t(d.list().size());
    }
    return MGO_OK;
}

// osecir::Simulation::advance(tmax) with dynamic NPIs over the reference tests' MockSimulation (state constant
// over time, cpp/tests/test_dynamic_npis.cpp:379-416) for node 0 of `model`, started at t0. Writes the
// effective contact matrix entry (0, 0) at each query time and the number of dampings of the matrix.
int32_t orc_npi_frozen_run(const mgo_graph_desc* model, double t0, double tmax, int32_t num_query,
                           const double* query_times, double* out_c00, int32_t* out_num_dampings)
{
    if (!valid_desc(model)) {
        return MGO_ERR_INVALID_VALUE;
    }
    OsecirModel m(model->num_groups);
    build_osecir(model, 0, m);
    mgo_integrator_cfg ic{MGO_INTEGRATOR_EULER, 0, 0, 0, 0};
    Simulation sim(m, make_core(ic), t0, 0.1);
    sim.frozen = true;
    sim.advance(tmax);
    auto* om = dynamic_cast<OsecirModel*>(sim.model.get());
    for (int q = 0; q < num_query; ++q) {
        out_c00[q] = om->contacts.at(query_times[q], 0, 0);
    }
    if (out_num_dampings) {
        *out_num_dampings = int32_t(om->contacts.mats[0].dampings.list().size());
    }
    return MGO_OK;
}

double orc_contact_at(const mgo_graph_desc* model, double t, int32_t i, int32_t j)
{
    OsecirModel m(model->num_groups);
    build_osecir(model, 0, m);
    return m.contacts.at(t, i, j);
}

} // extern "C"
