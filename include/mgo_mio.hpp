// mgo_mio.hpp - the mio:: call shapes of MEmilio's osecir Graph-ODE path over the C ABI of mgo.h.
//
// Header-only C++20 layer so that code written against
//   mio::osecir::Model / Parameters / ContactMatrixGroup,
//   mio::Graph<mio::SimulationNode<double, mio::osecir::Simulation<double>>, mio::MobilityEdge<double>>,
//   mio::make_mobility_sim(t0, dt, graph).advance(tmax), mio::osecir::simulate(...),
//   mio::ParameterStudy<...>::run(create_simulation, process_result)
// compiles against this engine with the same names, argument meanings and call order. The reference's
// own headers need Eigen and Boost; this layer has no dependencies: matrix and vector arguments may be
// any type with rows()/cols()/operator()(i, j) or size()/operator[](i) (Eigen types qualify), and small
// Matrix / Vector types with the Constant/Zero factories are provided for callers without Eigen.
//
// Reference interfaces mirrored (paths relative to the memilio source tree):
//   osecir parameters and model          cpp/models/ode_secir/parameters.h:46-394,400-450, model.h:56-80
//   populations                          cpp/memilio/epidemiology/populations.h (operator[], set_difference_*)
//   contact matrices and dampings        cpp/memilio/epidemiology/contact_matrix.h:190-194,392-400, damping.h
//   integrator cores                     cpp/memilio/math/stepper_wrapper.h:83-91, adapt_rk.h:152-175,
//                                        euler.h:32-63, compartments/simulation_base.h:86-89
//   Graph, add_node, add_edge            cpp/memilio/mobility/graph.h:144-303 (edges kept sorted by
//                                        (start, end), a repeated edge replaces the old one)
//   SimulationNode, MobilityEdge         cpp/memilio/mobility/metapopulation_mobility_instant.h:39-101,303-363
//   make_mobility_sim, GraphSimulation   metapopulation_mobility_instant.h:656-678, graph_simulation.h:60-120
//   osecir::simulate                     cpp/models/ode_secir/model.h:385-390
//   ParameterStudy::run / run_serial     cpp/memilio/compartments/parameter_studies.h:47-130,156-217,288-345
//   interpolate_simulation_result        cpp/memilio/data/analyze_result.h:186-265
//   interpolate_ensemble_results,        analyze_result.h:77-86, 267-343 (the statistics computed on the
//   sum_nodes, ensemble_mean,            device by mgo_sum_nodes_values, mgo_ensemble_*_values)
//   ensemble_percentile
//   parameter sampling                   cpp/memilio/utils/uncertain_value.h, parameter_distributions.h,
//                                        epidemiology/uncertain_matrix.h, damping_sampling.h:255-306,
//                                        models/ode_secir/parameter_space.h:46-251
//
// Differences a caller sees:
//   * Results are daily: a node's get_result() holds the values at the integer days reached (what
//     interpolate_simulation_result would return for the reference's step-wise series) plus the current
//     state when the simulation time is not an integer day.
//   * A node's ContactMatrixGroup may hold several matrices if at most one of them has dampings; they
//     are folded into one (exact in real arithmetic, ULP-level rounding differences).
//   * All nodes of a graph share one integrator core configuration and initial step (the device runs a
//     graph with one configuration); differing nodes raise Error(MGO_ERR_UNSUPPORTED).
//   * ParameterStudy::run executes every simulation create_simulation returns in one device batch, then
//     calls process_result in run order; create_simulation samples as in the reference (draw_sample,
//     make_sampled_graph_simulation).
//   * Parameter sampling (UncertainValue distributions, set_params_distributions_normal, draw_sample) follows
//     the reference's rules with std::mt19937_64 in place of its Random123 generator: samples follow the
//     same distributions, not the same values for a seed.
//   * Errors are thrown as mgo_mio::Error (status + message) instead of the reference's asserts.
//   * Edge result tracking (MobilityEdge save_indices / get_mobility_results) is not provided; passing save
//     indices is accepted and ignored.
//
// Parameter-graph JSON files (mio::write_graph / read_graph) are in mgo_mio_io.hpp.
//
// The names live in namespace mgo_mio; unless MGO_MIO_NO_ALIAS is defined, `mio` is an alias of it.
#ifndef MGO_MIO_HPP
#define MGO_MIO_HPP

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "mgo.h"

namespace mgo_mio
{

using ScalarType = double;

// ---------------------------------------------------------------------------------------------- errors
class Error : public std::runtime_error
{
public:
    Error(mgo_status st, const std::string& msg)
        : std::runtime_error("mgo status " + std::to_string(st) + ": " + msg)
        , status(st)
    {
    }
    mgo_status status;
};

namespace detail
{
inline void check(mgo_status st, const mgo_context* ctx)
{
    if (st != MGO_OK) {
        throw Error(st, mgo_last_error(ctx));
    }
}

struct ContextHolder {
    mgo_context* ctx = nullptr;
    mgo_device_cfg cfg{0, 0, 0};
    ~ContextHolder()
    {
        if (ctx) {
            mgo_context_destroy(ctx);
        }
    }
};

inline ContextHolder& context_holder()
{
    static ContextHolder h;
    return h;
}
} // namespace detail

// Device context used by every simulation of this layer (created on first use, device 0 by default).
inline mgo_context* default_context()
{
    auto& h = detail::context_holder();
    if (!h.ctx) {
        detail::check(mgo_context_create(&h.cfg, &h.ctx), nullptr);
    }
    return h.ctx;
}

// Selects the device (and HBM budget, MGO_CTX_* flags) for simulations created afterwards.
inline void set_device(int device, size_t hbm_budget_bytes = 0, int flags = 0)
{
    auto& h = detail::context_holder();
    if (h.ctx) {
        mgo_context_destroy(h.ctx);
        h.ctx = nullptr;
    }
    h.cfg = mgo_device_cfg{device, hbm_budget_bytes, flags};
}

// ---------------------------------------------------------------------------------------------- indices
class AgeGroup
{
public:
    constexpr explicit AgeGroup(size_t i = 0)
        : m_i(i)
    {
    }
    constexpr size_t get() const
    {
        return m_i;
    }
    AgeGroup& operator++()
    {
        ++m_i;
        return *this;
    }
    friend constexpr bool operator==(AgeGroup a, AgeGroup b)
    {
        return a.m_i == b.m_i;
    }
    friend constexpr auto operator<=>(AgeGroup a, AgeGroup b)
    {
        return a.m_i <=> b.m_i;
    }

private:
    size_t m_i;
};

struct SimulationTime {
    explicit SimulationTime(double t)
        : m_t(t)
    {
    }
    double get() const
    {
        return m_t;
    }
    double m_t;
};

struct DampingLevel {
    explicit DampingLevel(int v)
        : m_v(v)
    {
    }
    int get() const
    {
        return m_v;
    }
    int m_v;
};

struct DampingType {
    explicit DampingType(int v)
        : m_v(v)
    {
    }
    int get() const
    {
        return m_v;
    }
    int m_v;
};

// ---------------------------------------------------------------------------------------------- dense types
// Row-major dense matrix for callers without Eigen.
class Matrix
{
public:
    Matrix() = default;
    Matrix(size_t rows, size_t cols, double v = 0.0)
        : m_rows(rows)
        , m_cols(cols)
        , m_data(rows * cols, v)
    {
    }
    template <class M>
        requires requires(const M& m) {
            m.rows();
            m.cols();
            m(0, 0);
        }
    static Matrix from(const M& m)
    {
        Matrix r(size_t(m.rows()), size_t(m.cols()));
        for (size_t i = 0; i < r.m_rows; ++i) {
            for (size_t j = 0; j < r.m_cols; ++j) {
                r(i, j) = double(m(i, j));
            }
        }
        return r;
    }
    static Matrix Constant(size_t rows, size_t cols, double v)
    {
        return Matrix(rows, cols, v);
    }
    static Matrix Zero(size_t rows, size_t cols)
    {
        return Matrix(rows, cols, 0.0);
    }
    double& operator()(size_t i, size_t j)
    {
        return m_data[i * m_cols + j];
    }
    double operator()(size_t i, size_t j) const
    {
        return m_data[i * m_cols + j];
    }
    size_t rows() const
    {
        return m_rows;
    }
    size_t cols() const
    {
        return m_cols;
    }
    const double* data() const
    {
        return m_data.data();
    }
    friend bool operator==(const Matrix& a, const Matrix& b)
    {
        return a.m_rows == b.m_rows && a.m_cols == b.m_cols && a.m_data == b.m_data;
    }

private:
    size_t m_rows = 0, m_cols = 0;
    std::vector<double> m_data;
};

class Vector
{
public:
    Vector() = default;
    explicit Vector(size_t n, double v = 0.0)
        : m_data(n, v)
    {
    }
    template <class V>
        requires requires(const V& v) {
            v.size();
            v[0];
        }
    static Vector from(const V& v)
    {
        Vector r(size_t(v.size()));
        for (size_t i = 0; i < r.size(); ++i) {
            r[i] = double(v[i]);
        }
        return r;
    }
    static Vector Constant(size_t n, double v)
    {
        return Vector(n, v);
    }
    static Vector Zero(size_t n)
    {
        return Vector(n, 0.0);
    }
    size_t size() const
    {
        return m_data.size();
    }
    double& operator[](size_t i)
    {
        return m_data[i];
    }
    double operator[](size_t i) const
    {
        return m_data[i];
    }
    double& operator()(size_t i)
    {
        return m_data[i];
    }
    double operator()(size_t i) const
    {
        return m_data[i];
    }
    const double* data() const
    {
        return m_data.data();
    }
    double* data()
    {
        return m_data.data();
    }

private:
    std::vector<double> m_data;
};

namespace detail
{
template <class M>
Matrix to_matrix(const M& m)
{
    if constexpr (std::is_same_v<M, Matrix>) {
        return m;
    }
    else {
        return Matrix::from(m);
    }
}
template <class V>
Vector to_vector(const V& v)
{
    if constexpr (std::is_same_v<V, Vector>) {
        return v;
    }
    else if constexpr (requires { v.size(); v[0]; }) {
        return Vector::from(v);
    }
    else {
        // Eigen vectors index with operator()
        Vector r(size_t(v.size()));
        for (size_t i = 0; i < r.size(); ++i) {
            r[i] = double(v(i));
        }
        return r;
    }
}
} // namespace detail

// ---------------------------------------------------------------------------------------------- contacts
struct Damping {
    Matrix coeffs;
    int level = 0;
    int type  = 0;
    double time = 0.0;
};

// ContactMatrix: baseline B, minimum M, dampings; effective B - D(t) o (B - M).
class ContactMatrix
{
public:
    explicit ContactMatrix(size_t num_groups = 0)
        : m_baseline(num_groups, num_groups)
        , m_minimum(num_groups, num_groups)
    {
    }
    template <class M>
    explicit ContactMatrix(const M& baseline)
        : m_baseline(detail::to_matrix(baseline))
        , m_minimum(m_baseline.rows(), m_baseline.cols())
    {
    }
    template <class M1, class M2>
    ContactMatrix(const M1& baseline, const M2& minimum)
        : m_baseline(detail::to_matrix(baseline))
        , m_minimum(detail::to_matrix(minimum))
    {
    }
    template <class M>
    void add_damping(const M& coeffs, DampingLevel level, DampingType type, SimulationTime t)
    {
        m_dampings.push_back(Damping{detail::to_matrix(coeffs), level.get(), type.get(), t.get()});
    }
    template <class M>
    void add_damping(const M& coeffs, SimulationTime t)
    {
        add_damping(coeffs, DampingLevel(0), DampingType(0), t);
    }
    void add_damping(double d, DampingLevel level, DampingType type, SimulationTime t)
    {
        add_damping(Matrix::Constant(m_baseline.rows(), m_baseline.cols(), d), level, type, t);
    }
    void add_damping(double d, SimulationTime t)
    {
        add_damping(d, DampingLevel(0), DampingType(0), t);
    }
    void clear_dampings()
    {
        m_dampings.clear();
    }
    const Matrix& get_baseline() const
    {
        return m_baseline;
    }
    Matrix& get_baseline()
    {
        return m_baseline;
    }
    const Matrix& get_minimum() const
    {
        return m_minimum;
    }
    Matrix& get_minimum()
    {
        return m_minimum;
    }
    const std::vector<Damping>& get_dampings() const
    {
        return m_dampings;
    }

private:
    Matrix m_baseline, m_minimum;
    std::vector<Damping> m_dampings;
};

class ContactMatrixGroup
{
public:
    ContactMatrixGroup(size_t num_matrices = 1, size_t num_groups = 0)
        : m_mats(num_matrices, ContactMatrix(num_groups))
    {
    }
    ContactMatrix& operator[](size_t k)
    {
        return m_mats.at(k);
    }
    const ContactMatrix& operator[](size_t k) const
    {
        return m_mats.at(k);
    }
    size_t get_num_matrices() const
    {
        return m_mats.size();
    }
    template <class... Args>
    void add_damping(const Args&... args)
    {
        for (auto& m : m_mats) {
            m.add_damping(args...);
        }
    }
    void clear_dampings()
    {
        for (auto& m : m_mats) {
            m.clear_dampings();
        }
    }

private:
    std::vector<ContactMatrix> m_mats;
};

// ---------------------------------------------------------------------------------------------- results
// Time series of one node: num_elements values per time point.
template <class FP = double>
class TimeSeries
{
public:
    explicit TimeSeries(size_t num_elements = 0)
        : m_n(num_elements)
    {
    }
    size_t get_num_elements() const
    {
        return m_n;
    }
    size_t get_num_time_points() const
    {
        return m_t.size();
    }
    void add_time_point(FP t, const FP* values)
    {
        m_t.push_back(t);
        m_v.insert(m_v.end(), values, values + m_n);
    }
    template <class V>
        requires(!std::is_convertible_v<const V&, const FP*>)
    void add_time_point(FP t, const V& values)
    {
        const Vector v = detail::to_vector(values);
        add_time_point(t, v.data());
    }
    FP get_time(size_t i) const
    {
        return m_t.at(i);
    }
    FP get_last_time() const
    {
        return m_t.back();
    }
    Vector get_value(size_t i) const
    {
        Vector r(m_n);
        std::copy(m_v.begin() + std::ptrdiff_t(i * m_n), m_v.begin() + std::ptrdiff_t((i + 1) * m_n), r.data());
        return r;
    }
    Vector get_last_value() const
    {
        return get_value(m_t.size() - 1);
    }
    const FP* value_data(size_t i) const
    {
        return m_v.data() + i * m_n;
    }
    void clear()
    {
        m_t.clear();
        m_v.clear();
    }

private:
    size_t m_n;
    std::vector<FP> m_t;
    std::vector<FP> m_v;
};

// Values at the integer days ceil(t0 - abs_tol) .. floor(tN + abs_tol), linearly interpolated between the
// series' time points (analyze_result.h:186-265).
template <class FP>
TimeSeries<FP> interpolate_simulation_result(const TimeSeries<FP>& ts, FP abs_tol = 1e-10)
{
    TimeSeries<FP> out(ts.get_num_elements());
    const size_t np = ts.get_num_time_points();
    if (np == 0) {
        return out;
    }
    const long d0 = long(std::ceil(ts.get_time(0) - abs_tol));
    const long d1 = long(std::floor(ts.get_last_time() + abs_tol));
    std::vector<FP> v(ts.get_num_elements());
    size_t k = 0;
    for (long d = d0; d <= d1; ++d) {
        const FP day = FP(d);
        while (k + 1 < np && ts.get_time(k + 1) < day - abs_tol) {
            ++k;
        }
        if (std::abs(ts.get_time(k) - day) <= abs_tol || k + 1 >= np) {
            out.add_time_point(day, ts.value_data(k));
            continue;
        }
        if (std::abs(ts.get_time(k + 1) - day) <= abs_tol) {
            out.add_time_point(day, ts.value_data(k + 1));
            continue;
        }
        const FP w = (day - ts.get_time(k)) / (ts.get_time(k + 1) - ts.get_time(k));
        for (size_t i = 0; i < v.size(); ++i) {
            v[i] = ts.value_data(k)[i] + w * (ts.value_data(k + 1)[i] - ts.value_data(k)[i]);
        }
        out.add_time_point(day, v.data());
    }
    return out;
}

// ---------------------------------------------------------------------------------------------- integrators
// Integrator cores; each names the device integrator it selects (include/mgo.h MGO_INTEGRATOR_*).
template <class FP = double>
class OdeIntegratorCore
{
public:
    virtual ~OdeIntegratorCore()                                 = default;
    virtual mgo_integrator_cfg cfg() const                       = 0;
    virtual std::unique_ptr<OdeIntegratorCore<FP>> clone() const = 0;
};

namespace detail
{
template <class FP, int Kind>
class AdaptiveCore : public OdeIntegratorCore<FP>
{
public:
    AdaptiveCore(FP abs_tol = 1e-10, FP rel_tol = 1e-5, FP dt_min = std::numeric_limits<FP>::min(),
                 FP dt_max = std::numeric_limits<FP>::max())
        : m_abs(abs_tol)
        , m_rel(rel_tol)
        , m_min(dt_min)
        , m_max(dt_max)
    {
    }
    void set_abs_tolerance(FP t)
    {
        m_abs = t;
    }
    void set_rel_tolerance(FP t)
    {
        m_rel = t;
    }
    void set_dt_min(FP t)
    {
        m_min = t;
    }
    void set_dt_max(FP t)
    {
        m_max = t;
    }
    FP get_dt_min() const
    {
        return m_min;
    }
    FP get_dt_max() const
    {
        return m_max;
    }
    mgo_integrator_cfg cfg() const override
    {
        return mgo_integrator_cfg{Kind, double(m_abs), double(m_rel), double(m_min), double(m_max)};
    }

protected:
    FP m_abs, m_rel, m_min, m_max;
};
} // namespace detail

// ControlledStepperWrapper<FP, runge_kutta_cash_karp54>: the reference's default core (odeint controller)
template <class FP = double>
class ControlledStepperWrapper : public detail::AdaptiveCore<FP, MGO_INTEGRATOR_CASH_KARP54>
{
public:
    using detail::AdaptiveCore<FP, MGO_INTEGRATOR_CASH_KARP54>::AdaptiveCore;
    std::unique_ptr<OdeIntegratorCore<FP>> clone() const override
    {
        return std::make_unique<ControlledStepperWrapper>(*this);
    }
};

// RKIntegratorCore with the default Runge-Kutta-Fehlberg 4(5) tableau
template <class FP = double>
class RKIntegratorCore : public detail::AdaptiveCore<FP, MGO_INTEGRATOR_RKF45>
{
public:
    using detail::AdaptiveCore<FP, MGO_INTEGRATOR_RKF45>::AdaptiveCore;
    std::unique_ptr<OdeIntegratorCore<FP>> clone() const override
    {
        return std::make_unique<RKIntegratorCore>(*this);
    }
};

// ExplicitStepperWrapper<FP, runge_kutta_cash_karp54>: fixed step (the simulation's dt)
template <class FP = double>
class ExplicitStepperWrapper : public OdeIntegratorCore<FP>
{
public:
    mgo_integrator_cfg cfg() const override
    {
        return mgo_integrator_cfg{MGO_INTEGRATOR_CASH_KARP54_FIXED, 0.0, 0.0, 0.0, 0.0};
    }
    std::unique_ptr<OdeIntegratorCore<FP>> clone() const override
    {
        return std::make_unique<ExplicitStepperWrapper>(*this);
    }
};

template <class FP = double>
class EulerIntegratorCore : public OdeIntegratorCore<FP>
{
public:
    mgo_integrator_cfg cfg() const override
    {
        return mgo_integrator_cfg{MGO_INTEGRATOR_EULER, 0.0, 0.0, 0.0, 0.0};
    }
    std::unique_ptr<OdeIntegratorCore<FP>> clone() const override
    {
        return std::make_unique<EulerIntegratorCore>(*this);
    }
};

// ---------------------------------------------------------------------------------------------- osecir
// DampingSampling (cpp/memilio/epidemiology/damping_sampling.h:40-230): one damping a dynamic NPI adds,
// value * group weights on the listed matrices of the contact matrix group.
// ---------------------------------------------------------------------------------------------- sampling
// Random numbers for parameter sampling (mio::thread_local_rng, utils/random_number_generator.h). The
// reference draws from a Random123 counter-based generator; this layer uses std::mt19937_64 with the
// standard library's distributions, so samples follow the reference's distributions, not its values.
class RandomNumberGenerator
{
public:
    using result_type = std::mt19937_64::result_type;
    RandomNumberGenerator()
    {
        std::random_device rd;
        seed({rd(), rd(), rd(), rd()});
    }
    static constexpr result_type min()
    {
        return std::mt19937_64::min();
    }
    static constexpr result_type max()
    {
        return std::mt19937_64::max();
    }
    result_type operator()()
    {
        return m_gen();
    }
    void seed(const std::vector<uint32_t>& seeds)
    {
        std::seed_seq sq(seeds.begin(), seeds.end());
        m_gen.seed(sq);
    }

private:
    std::mt19937_64 m_gen;
};

inline RandomNumberGenerator& thread_local_rng()
{
    thread_local RandomNumberGenerator rng;
    return rng;
}

// ParameterDistribution (utils/parameter_distributions.h:72-130): predefined samples are returned first, in
// the order added, each once; then random samples.
class ParameterDistribution
{
public:
    virtual ~ParameterDistribution() = default;
    void add_predefined_sample(double v)
    {
        m_predefined.push_back(v);
    }
    void remove_predefined_samples()
    {
        m_predefined.clear();
    }
    const std::vector<double>& get_predefined_samples() const
    {
        return m_predefined;
    }
    double get_sample(RandomNumberGenerator& rng)
    {
        if (!m_predefined.empty()) {
            const double v = m_predefined.front();
            m_predefined.erase(m_predefined.begin());
            return v;
        }
        return get_rand_sample(rng);
    }
    virtual double get_rand_sample(RandomNumberGenerator& rng) = 0;
    virtual std::unique_ptr<ParameterDistribution> clone() const = 0;

protected:
    std::vector<double> m_predefined;
};

// ParameterDistributionNormal (parameter_distributions.h:159-427): a normal distribution truncated to
// [lb, ub] (default [DBL_MIN, DBL_MAX]). Before drawing, a mean outside the bounds becomes (ub - lb) / 2 and
// the standard deviation shrinks until mean +- quantile * std fits; a draw outside the bounds is redrawn up
// to 10 times, then set to ub if the last draw is above it, else to lb.
class ParameterDistributionNormal : public ParameterDistribution
{
public:
    ParameterDistributionNormal() = default;
    ParameterDistributionNormal(double mean, double standard_dev)
        : m_mean(mean)
        , m_std(standard_dev)
    {
    }
    // standard deviation: the widest for which mean +- quantile * std fits in [lb, ub]
    ParameterDistributionNormal(double lower_bound, double upper_bound, double mean, double quantile)
        : m_mean(mean)
        , m_std(upper_bound)
        , m_lb(lower_bound)
        , m_ub(upper_bound)
        , m_quantile(quantile)
    {
        adapt_standard_dev();
    }
    ParameterDistributionNormal(double lower_bound, double upper_bound, double mean, double standard_dev,
                                double quantile)
        : m_mean(mean)
        , m_std(standard_dev)
        , m_lb(lower_bound)
        , m_ub(upper_bound)
        , m_quantile(quantile)
    {
        check_quantiles();
    }
    double get_mean() const
    {
        return m_mean;
    }
    void set_mean(double v)
    {
        m_mean = v;
    }
    double get_standard_dev() const
    {
        return m_std;
    }
    void set_standard_dev(double v)
    {
        m_std = v;
    }
    double get_lower_bound() const
    {
        return m_lb;
    }
    void set_lower_bound(double v)
    {
        m_lb = v;
    }
    double get_upper_bound() const
    {
        return m_ub;
    }
    void set_upper_bound(double v)
    {
        m_ub = v;
    }
    double get_quantile() const
    {
        return m_quantile;
    }
    bool check_quantiles()
    {
        bool changed = false;
        if (m_mean < m_lb || m_mean > m_ub) {
            m_mean  = 0.5 * (m_ub - m_lb);
            changed = true;
        }
        return adapt_standard_dev() || changed;
    }
    double get_rand_sample(RandomNumberGenerator& rng) override
    {
        if (m_ub == m_lb) {
            return m_lb;
        }
        check_quantiles();
        std::normal_distribution<double> nd(m_mean, m_std);
        double r = nd(rng);
        for (int i = 0; (r > m_ub || r < m_lb) && i < 10;) {
            r = nd(rng);
            if (++i == 10) {
                r = r > m_ub ? m_ub : m_lb;
            }
        }
        return r;
    }
    std::unique_ptr<ParameterDistribution> clone() const override
    {
        return std::make_unique<ParameterDistributionNormal>(*this);
    }

private:
    bool adapt_standard_dev()
    {
        bool changed = false;
        if (m_mean + m_std * m_quantile > m_ub) {
            m_std   = (m_ub - m_mean) / m_quantile;
            changed = true;
        }
        if (m_mean - m_std * m_quantile < m_lb) {
            m_std   = (m_mean - m_lb) / m_quantile;
            changed = true;
        }
        return changed;
    }
    double m_mean = 0.0;
    double m_std  = 1.0;
    double m_lb   = DBL_MIN;
    double m_ub   = DBL_MAX;
    double m_quantile = 2.5758;
};

// ParameterDistributionUniform (parameter_distributions.h:439-560)
class ParameterDistributionUniform : public ParameterDistribution
{
public:
    ParameterDistributionUniform(double lower_bound = 0.0, double upper_bound = 1.0)
        : m_lb(lower_bound)
        , m_ub(upper_bound)
    {
    }
    double get_lower_bound() const
    {
        return m_lb;
    }
    double get_upper_bound() const
    {
        return m_ub;
    }
    void set_lower_bound(double v)
    {
        m_lb = v;
    }
    void set_upper_bound(double v)
    {
        m_ub = v;
    }
    double get_rand_sample(RandomNumberGenerator& rng) override
    {
        return std::uniform_real_distribution<double>(m_lb, m_ub)(rng);
    }
    std::unique_ptr<ParameterDistribution> clone() const override
    {
        return std::make_unique<ParameterDistributionUniform>(*this);
    }

private:
    double m_lb, m_ub;
};

// ParameterDistributionLogNormal (parameter_distributions.h:566-690): exp of a normal (log_mean, log_stddev)
class ParameterDistributionLogNormal : public ParameterDistribution
{
public:
    ParameterDistributionLogNormal(double log_mean = 0.0, double log_stddev = 1.0)
        : m_mean(log_mean)
        , m_std(log_stddev)
    {
    }
    double get_log_mean() const
    {
        return m_mean;
    }
    double get_log_stddev() const
    {
        return m_std;
    }
    double get_rand_sample(RandomNumberGenerator& rng) override
    {
        return std::lognormal_distribution<double>(m_mean, m_std)(rng);
    }
    std::unique_ptr<ParameterDistribution> clone() const override
    {
        return std::make_unique<ParameterDistributionLogNormal>(*this);
    }

private:
    double m_mean, m_std;
};

// ParameterDistributionExponential (parameter_distributions.h:693-808)
class ParameterDistributionExponential : public ParameterDistribution
{
public:
    explicit ParameterDistributionExponential(double rate = 1.0)
        : m_rate(rate)
    {
    }
    double get_rate() const
    {
        return m_rate;
    }
    double get_rand_sample(RandomNumberGenerator& rng) override
    {
        return std::exponential_distribution<double>(m_rate)(rng);
    }
    std::unique_ptr<ParameterDistribution> clone() const override
    {
        return std::make_unique<ParameterDistributionExponential>(*this);
    }

private:
    double m_rate;
};

// ParameterDistributionConstant (parameter_distributions.h:812-920)
class ParameterDistributionConstant : public ParameterDistribution
{
public:
    explicit ParameterDistributionConstant(double constant = 0.0)
        : m_c(constant)
    {
    }
    double get_constant() const
    {
        return m_c;
    }
    double get_rand_sample(RandomNumberGenerator&) override
    {
        return m_c;
    }
    std::unique_ptr<ParameterDistribution> clone() const override
    {
        return std::make_unique<ParameterDistributionConstant>(*this);
    }

private:
    double m_c;
};

// non-owning pointer to a value's distribution (the reference's observer_ptr)
template <class T>
class observer_ptr
{
public:
    explicit observer_ptr(T* p = nullptr)
        : m_p(p)
    {
    }
    T* get() const
    {
        return m_p;
    }
    T* operator->() const
    {
        return m_p;
    }
    T& operator*() const
    {
        return *m_p;
    }
    explicit operator bool() const
    {
        return m_p != nullptr;
    }

private:
    T* m_p;
};

// UncertainValue (utils/uncertain_value.h:44-190): a value and an optional distribution it can be drawn
// from. Converts to the value (FP&), so it is read and assigned like a number; assigning a number keeps
// the distribution, copying copies it.
template <class FP = double>
class UncertainValue
{
public:
    UncertainValue(FP v = FP(0))
        : m_value(v)
    {
    }
    UncertainValue(const UncertainValue& o)
        : m_value(o.m_value)
        , m_dist(o.m_dist ? o.m_dist->clone() : nullptr)
    {
    }
    UncertainValue(UncertainValue&&) noexcept = default;
    UncertainValue& operator=(const UncertainValue& o)
    {
        if (this != &o) {
            m_value = o.m_value;
            m_dist  = o.m_dist ? o.m_dist->clone() : nullptr;
        }
        return *this;
    }
    UncertainValue& operator=(UncertainValue&&) noexcept = default;
    UncertainValue& operator=(FP v)
    {
        m_value = v;
        return *this;
    }
    operator FP&()
    {
        return m_value;
    }
    operator const FP&() const
    {
        return m_value;
    }
    FP& value()
    {
        return m_value;
    }
    const FP& value() const
    {
        return m_value;
    }
    void set_distribution(const ParameterDistribution& d)
    {
        m_dist = d.clone();
    }
    observer_ptr<ParameterDistribution> get_distribution()
    {
        return observer_ptr<ParameterDistribution>(m_dist.get());
    }
    observer_ptr<const ParameterDistribution> get_distribution() const
    {
        return observer_ptr<const ParameterDistribution>(m_dist.get());
    }
    // value = a sample of the distribution (unchanged without one)
    FP draw_sample()
    {
        if (m_dist) {
            m_value = m_dist->get_sample(thread_local_rng());
        }
        return m_value;
    }

private:
    FP m_value;
    std::unique_ptr<ParameterDistribution> m_dist;
};

// DampingSampling (epidemiology/damping_sampling.h:40-230): a damping of value * group weights that a
// dynamic NPI or a parameter study adds to the listed contact matrices; the value may carry a distribution.
template <class FP = double>
class DampingSampling
{
public:
    template <class V>
    DampingSampling(const UncertainValue<FP>& value, DampingLevel level, DampingType type, SimulationTime time,
                    std::vector<size_t> matrix_indices, const V& group_weights)
        : m_value(value)
        , m_level(level)
        , m_type(type)
        , m_time(time)
        , m_mats(std::move(matrix_indices))
        , m_weights(detail::to_vector(group_weights))
    {
    }
    const UncertainValue<FP>& get_value() const
    {
        return m_value;
    }
    UncertainValue<FP>& get_value()
    {
        return m_value;
    }
    DampingLevel get_level() const
    {
        return m_level;
    }
    void set_level(DampingLevel l)
    {
        m_level = l;
    }
    DampingType get_type() const
    {
        return m_type;
    }
    void set_type(DampingType t)
    {
        m_type = t;
    }
    SimulationTime get_time() const
    {
        return m_time;
    }
    void set_time(SimulationTime t)
    {
        m_time = t;
    }
    const std::vector<size_t>& get_matrix_indices() const
    {
        return m_mats;
    }
    const Vector& get_group_weights() const
    {
        return m_weights;
    }
    void draw_sample()
    {
        m_value.draw_sample();
    }

private:
    UncertainValue<FP> m_value;
    DampingLevel m_level;
    DampingType m_type;
    SimulationTime m_time;
    std::vector<size_t> m_mats;
    Vector m_weights;
};

// make_contact_damping_matrix (damping_sampling.h:268-290): d_ij = 1 - sqrt((1 - g_i)(1 - g_j))
inline Matrix make_contact_damping_matrix(const Vector& g)
{
    Matrix m(g.size(), g.size());
    for (size_t i = 0; i < g.size(); ++i) {
        for (size_t j = 0; j < g.size(); ++j) {
            m(i, j) = 1.0 - std::sqrt((1.0 - g[i]) * (1.0 - g[j]));
        }
    }
    return m;
}

// make_mobility_damping_vector (damping_sampling.h:298-306): each group's value on all its compartments
inline Vector make_mobility_damping_vector(size_t size, const Vector& g)
{
    Vector v(size);
    const size_t per = g.size() ? size / g.size() : 0;
    for (size_t i = 0; i < size; ++i) {
        v[i] = per ? g[i / per] : 0.0;
    }
    return v;
}

// apply_dampings (damping_sampling.h:255-266): every damping sampling adds make_matrix(value * weights) to
// each matrix it lists
template <class Group, class DampingRange, class F>
void apply_dampings(Group& group, const DampingRange& dampings, F make_matrix)
{
    for (const auto& d : dampings) {
        Vector w = d.get_group_weights();
        for (size_t i = 0; i < w.size(); ++i) {
            w[i] *= double(d.get_value());
        }
        for (size_t k : d.get_matrix_indices()) {
            group[k].add_damping(make_matrix(w), d.get_level(), d.get_type(), d.get_time());
        }
    }
}

// DynamicNPIs (cpp/memilio/epidemiology/dynamic_npis.h:36-242): thresholds (cases per base value, sorted
// descending, an equal threshold replaced) and the NPIs they trigger; duration, implementation delay and
// the directive interval in which NPIs can start.
template <class FP = double>
class DynamicNPIs
{
public:
    void set_threshold(double threshold, const std::vector<DampingSampling<FP>>& dampings)
    {
        auto it = m_thresholds.begin();
        while (it != m_thresholds.end() && it->first > threshold) {
            ++it;
        }
        if (it != m_thresholds.end() && !(threshold > it->first)) {
            it->second = dampings;
        }
        else {
            m_thresholds.insert(it, std::make_pair(threshold, dampings));
        }
    }
    const std::vector<std::pair<double, std::vector<DampingSampling<FP>>>>& get_thresholds() const
    {
        return m_thresholds;
    }
    SimulationTime get_duration() const
    {
        return m_duration;
    }
    void set_duration(SimulationTime v)
    {
        m_duration = v;
    }
    SimulationTime get_implementation_delay() const
    {
        return m_delay;
    }
    void set_implementation_delay(SimulationTime v)
    {
        m_delay = v;
    }
    double get_base_value() const
    {
        return m_base;
    }
    void set_base_value(double v)
    {
        m_base = v;
    }
    SimulationTime get_directive_begin() const
    {
        return m_begin;
    }
    void set_directive_begin(SimulationTime v)
    {
        m_begin = v;
    }
    SimulationTime get_directive_end() const
    {
        return m_end;
    }
    void set_directive_end(SimulationTime v)
    {
        m_end = v;
    }
    // samples the value of every damping of every threshold (dynamic_npis.h:192-199)
    void draw_sample()
    {
        for (auto& t : m_thresholds) {
            for (auto& d : t.second) {
                d.draw_sample();
            }
        }
    }

private:
    std::vector<std::pair<double, std::vector<DampingSampling<FP>>>> m_thresholds;
    SimulationTime m_duration{14.0};
    SimulationTime m_delay{0.0};
    SimulationTime m_begin{std::numeric_limits<double>::lowest()};
    SimulationTime m_end{std::numeric_limits<double>::max()};
    double m_base = 1.0;
};

namespace osecir
{

enum class InfectionState
{
    Susceptible = 0,
    Exposed,
    InfectedNoSymptoms,
    InfectedNoSymptomsConfirmed,
    InfectedSymptoms,
    InfectedSymptomsConfirmed,
    InfectedSevere,
    InfectedCritical,
    Recovered,
    Dead,
    Count
};

// Per-age-group parameter values (CustomIndexArray<UncertainValue, AgeGroup>); assigning a scalar sets
// every group (and, as assigning an UncertainValue without distribution, drops their distributions).
class AgeArray
{
public:
    explicit AgeArray(size_t num_groups = 0, double v = 0.0)
        : m_v(num_groups, UncertainValue<double>(v))
    {
    }
    AgeArray& operator=(double v)
    {
        std::fill(m_v.begin(), m_v.end(), UncertainValue<double>(v));
        return *this;
    }
    UncertainValue<double>& operator[](AgeGroup g)
    {
        return m_v.at(g.get());
    }
    const UncertainValue<double>& operator[](AgeGroup g) const
    {
        return m_v.at(g.get());
    }
    size_t numel() const
    {
        return m_v.size();
    }

private:
    std::vector<UncertainValue<double>> m_v;
};

// parameter tags (parameters.h:46-394): scalar ones index MGO_P_*, per-age ones MGO_A_*
#define MGO_MIO_SCALAR_TAG(Name, Index)                                                                         \
    template <class FP = double>                                                                                \
    struct Name {                                                                                               \
        static constexpr int scalar_index = Index;                                                              \
        static std::string name()                                                                               \
        {                                                                                                       \
            return #Name;                                                                                       \
        }                                                                                                       \
    };
#define MGO_MIO_AGE_TAG(Name, Index)                                                                            \
    template <class FP = double>                                                                                \
    struct Name {                                                                                               \
        static constexpr int age_index = Index;                                                                 \
        static std::string name()                                                                               \
        {                                                                                                       \
            return #Name;                                                                                       \
        }                                                                                                       \
    };
MGO_MIO_SCALAR_TAG(StartDay, MGO_P_START_DAY)
MGO_MIO_SCALAR_TAG(Seasonality, MGO_P_SEASONALITY)
MGO_MIO_SCALAR_TAG(ICUCapacity, MGO_P_ICU_CAPACITY)
MGO_MIO_SCALAR_TAG(TestAndTraceCapacity, MGO_P_TEST_AND_TRACE_CAPACITY)
MGO_MIO_SCALAR_TAG(TestAndTraceCapacityMaxRisk, MGO_P_TEST_AND_TRACE_MAX_RISK)
MGO_MIO_AGE_TAG(TimeExposed, MGO_A_TIME_EXPOSED)
MGO_MIO_AGE_TAG(TimeInfectedNoSymptoms, MGO_A_TIME_INFECTED_NO_SYMPTOMS)
MGO_MIO_AGE_TAG(TimeInfectedSymptoms, MGO_A_TIME_INFECTED_SYMPTOMS)
MGO_MIO_AGE_TAG(TimeInfectedSevere, MGO_A_TIME_INFECTED_SEVERE)
MGO_MIO_AGE_TAG(TimeInfectedCritical, MGO_A_TIME_INFECTED_CRITICAL)
MGO_MIO_AGE_TAG(TransmissionProbabilityOnContact, MGO_A_TRANSMISSION_PROBABILITY_ON_CONTACT)
MGO_MIO_AGE_TAG(RelativeTransmissionNoSymptoms, MGO_A_RELATIVE_TRANSMISSION_NO_SYMPTOMS)
MGO_MIO_AGE_TAG(RecoveredPerInfectedNoSymptoms, MGO_A_RECOVERED_PER_INFECTED_NO_SYMPTOMS)
MGO_MIO_AGE_TAG(RiskOfInfectionFromSymptomatic, MGO_A_RISK_OF_INFECTION_FROM_SYMPTOMATIC)
MGO_MIO_AGE_TAG(MaxRiskOfInfectionFromSymptomatic, MGO_A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC)
MGO_MIO_AGE_TAG(SeverePerInfectedSymptoms, MGO_A_SEVERE_PER_INFECTED_SYMPTOMS)
MGO_MIO_AGE_TAG(CriticalPerSevere, MGO_A_CRITICAL_PER_SEVERE)
MGO_MIO_AGE_TAG(DeathsPerSevere, MGO_A_DEATHS_PER_SEVERE)
MGO_MIO_AGE_TAG(DeathsPerCritical, MGO_A_DEATHS_PER_CRITICAL)
#undef MGO_MIO_SCALAR_TAG
#undef MGO_MIO_AGE_TAG

template <class FP = double>
struct ContactPatterns {
    static std::string name()
    {
        return "ContactPatterns";
    }
};

} // namespace osecir

// UncertainContactMatrix (epidemiology/uncertain_matrix.h:40-230): the contact matrix group and the
// parameter-study inputs beside it: contact dampings to sample, the school-holiday damping and the
// holidays. Converts to the ContactMatrixGroup.
template <class FP = double>
class UncertainContactMatrix
{
public:
    UncertainContactMatrix(size_t num_matrices = 1, size_t num_groups = 1)
        : UncertainContactMatrix(ContactMatrixGroup(num_matrices, num_groups))
    {
    }
    UncertainContactMatrix(const ContactMatrixGroup& cont_freq)
        : m_cont_freq(cont_freq)
        , m_holiday_damping(UncertainValue<FP>(0.0), DampingLevel(0), DampingType(0), SimulationTime(0.0), {},
                            Vector(cont_freq.get_num_matrices() ? cont_freq[0].get_baseline().rows() : 0, 0.0))
    {
    }
    // replaces the matrices; the sampling inputs stay
    UncertainContactMatrix& operator=(const ContactMatrixGroup& cont_freq)
    {
        m_cont_freq = cont_freq;
        return *this;
    }
    operator ContactMatrixGroup&()
    {
        return m_cont_freq;
    }
    operator const ContactMatrixGroup&() const
    {
        return m_cont_freq;
    }
    ContactMatrixGroup& get_cont_freq_mat()
    {
        return m_cont_freq;
    }
    const ContactMatrixGroup& get_cont_freq_mat() const
    {
        return m_cont_freq;
    }
    std::vector<DampingSampling<FP>>& get_dampings()
    {
        return m_dampings;
    }
    const std::vector<DampingSampling<FP>>& get_dampings() const
    {
        return m_dampings;
    }
    DampingSampling<FP>& get_school_holiday_damping()
    {
        return m_holiday_damping;
    }
    const DampingSampling<FP>& get_school_holiday_damping() const
    {
        return m_holiday_damping;
    }
    std::vector<std::pair<SimulationTime, SimulationTime>>& get_school_holidays()
    {
        return m_holidays;
    }
    const std::vector<std::pair<SimulationTime, SimulationTime>>& get_school_holidays() const
    {
        return m_holidays;
    }
    void draw_sample_dampings()
    {
        for (auto& d : m_dampings) {
            d.draw_sample();
        }
        m_holiday_damping.draw_sample();
    }
    // the matrices with the sampled dampings (uncertain_matrix.h:172-198): without accum the matrices'
    // dampings are replaced; each holiday switches the holiday damping on at its first day and to 0 at its
    // last
    ContactMatrixGroup make_matrix(bool accum = false)
    {
        if (!accum) {
            m_cont_freq.clear_dampings();
        }
        auto mk = [](const Vector& v) {
            return make_contact_damping_matrix(v);
        };
        apply_dampings(m_cont_freq, m_dampings, mk);
        for (const auto& h : m_holidays) {
            DampingSampling<FP> d = m_holiday_damping;
            d.set_time(h.first);
            apply_dampings(m_cont_freq, std::vector<DampingSampling<FP>>{d}, mk);
            d.get_value() = 0.0;
            d.set_time(h.second);
            apply_dampings(m_cont_freq, std::vector<DampingSampling<FP>>{d}, mk);
        }
        return m_cont_freq;
    }
    ContactMatrixGroup draw_sample(bool accum = false)
    {
        draw_sample_dampings();
        return make_matrix(accum);
    }

private:
    ContactMatrixGroup m_cont_freq;
    std::vector<DampingSampling<FP>> m_dampings;
    DampingSampling<FP> m_holiday_damping;
    std::vector<std::pair<SimulationTime, SimulationTime>> m_holidays;
};

namespace osecir
{

// DynamicNPIsInfectedSymptoms (parameters.h:342-355)
template <class FP = double>
struct DynamicNPIsInfectedSymptoms {
    static std::string name()
    {
        return "DynamicNPIsInfectedSymptoms";
    }
};

template <class FP = double>
class Parameters
{
public:
    explicit Parameters(AgeGroup num_groups)
        : m_G(num_groups.get())
        , m_contacts(1, num_groups.get())
    {
        m_scalars.fill(UncertainValue<double>(0.0));
        m_scalars[MGO_P_ICU_CAPACITY]            = std::numeric_limits<double>::max();
        m_scalars[MGO_P_TEST_AND_TRACE_CAPACITY] = std::numeric_limits<double>::max();
        m_scalars[MGO_P_TEST_AND_TRACE_MAX_RISK] = 5.0;
        m_ages.assign(MGO_NUM_AGE_PARAMS, AgeArray(m_G, 0.0));
        for (int k : {MGO_A_TIME_EXPOSED, MGO_A_TIME_INFECTED_NO_SYMPTOMS, MGO_A_TIME_INFECTED_SYMPTOMS,
                      MGO_A_TIME_INFECTED_SEVERE, MGO_A_TIME_INFECTED_CRITICAL,
                      MGO_A_TRANSMISSION_PROBABILITY_ON_CONTACT, MGO_A_RELATIVE_TRANSMISSION_NO_SYMPTOMS,
                      MGO_A_RISK_OF_INFECTION_FROM_SYMPTOMATIC, MGO_A_MAX_RISK_OF_INFECTION_FROM_SYMPTOMATIC}) {
            m_ages[size_t(k)] = 1.0;
        }
    }
    AgeGroup get_num_groups() const
    {
        return AgeGroup(m_G);
    }
    template <class Tag>
    decltype(auto) get()
    {
        if constexpr (requires { Tag::scalar_index; }) {
            return (m_scalars[Tag::scalar_index]);
        }
        else if constexpr (requires { Tag::age_index; }) {
            return (m_ages[Tag::age_index]);
        }
        else if constexpr (std::is_same_v<Tag, DynamicNPIsInfectedSymptoms<FP>>) {
            return (m_npis);
        }
        else {
            static_assert(std::is_same_v<Tag, ContactPatterns<FP>>, "parameter not on the device path");
            return (m_contacts);
        }
    }
    template <class Tag>
    decltype(auto) get() const
    {
        if constexpr (requires { Tag::scalar_index; }) {
            return (std::as_const(m_scalars[Tag::scalar_index]));
        }
        else if constexpr (requires { Tag::age_index; }) {
            return (std::as_const(m_ages[Tag::age_index]));
        }
        else if constexpr (std::is_same_v<Tag, DynamicNPIsInfectedSymptoms<FP>>) {
            return (std::as_const(m_npis));
        }
        else {
            static_assert(std::is_same_v<Tag, ContactPatterns<FP>>, "parameter not on the device path");
            return (std::as_const(m_contacts));
        }
    }
    template <class Tag, class V>
    void set(const V& v)
    {
        get<Tag>() = v;
    }
    FP& get_commuter_nondetection()
    {
        return m_scalars[MGO_P_COMMUTER_NONDETECTION];
    }
    FP get_commuter_nondetection() const
    {
        return m_scalars[MGO_P_COMMUTER_NONDETECTION];
    }
    FP& get_start_commuter_detection()
    {
        return m_scalars[MGO_P_START_COMMUTER_DETECTION];
    }
    FP get_start_commuter_detection() const
    {
        return m_scalars[MGO_P_START_COMMUTER_DETECTION];
    }
    FP& get_end_commuter_detection()
    {
        return m_scalars[MGO_P_END_COMMUTER_DETECTION];
    }
    FP get_end_commuter_detection() const
    {
        return m_scalars[MGO_P_END_COMMUTER_DETECTION];
    }
    // Clamps parameters to admissible values (parameters.h:465-612): seasonality outside [0, 0.5] -> 0,
    // negative capacities and TestAndTraceCapacityMaxRisk -> 0, stay times below 0.1 -> 0.1,
    // RelativeTransmissionNoSymptoms < 0 -> 0, the other probabilities outside [0, 1] -> 0 (not
    // MaxRiskOfInfectionFromSymptomatic), DeathsPerSevere -> 0 where CriticalPerSevere + DeathsPerSevere > 1.
    // Returns true if anything was corrected.
    bool apply_constraints()
    {
        bool corrected = false;
        auto fix       = [&](UncertainValue<double>& v, bool bad, double to) {
            if (bad) {
                v         = to;
                corrected = true;
            }
        };
        auto& sea = m_scalars[MGO_P_SEASONALITY];
        fix(sea, sea < 0.0 || sea > 0.5, 0.0);
        for (int k : {MGO_P_ICU_CAPACITY, MGO_P_TEST_AND_TRACE_CAPACITY, MGO_P_TEST_AND_TRACE_MAX_RISK}) {
            fix(m_scalars[size_t(k)], m_scalars[size_t(k)] < 0.0, 0.0);
        }
        for (size_t g = 0; g < m_G; ++g) {
            const AgeGroup a(g);
            for (int k : {MGO_A_TIME_EXPOSED, MGO_A_TIME_INFECTED_NO_SYMPTOMS, MGO_A_TIME_INFECTED_SYMPTOMS,
                          MGO_A_TIME_INFECTED_SEVERE, MGO_A_TIME_INFECTED_CRITICAL}) {
                auto& v = m_ages[size_t(k)][a];
                fix(v, v < 0.1, 0.1);
            }
            auto& rel = m_ages[MGO_A_RELATIVE_TRANSMISSION_NO_SYMPTOMS][a];
            fix(rel, rel < 0.0, 0.0);
            for (int k : {MGO_A_TRANSMISSION_PROBABILITY_ON_CONTACT, MGO_A_RECOVERED_PER_INFECTED_NO_SYMPTOMS,
                          MGO_A_RISK_OF_INFECTION_FROM_SYMPTOMATIC, MGO_A_SEVERE_PER_INFECTED_SYMPTOMS,
                          MGO_A_CRITICAL_PER_SEVERE, MGO_A_DEATHS_PER_SEVERE}) {
                auto& v = m_ages[size_t(k)][a];
                fix(v, v < 0.0 || v > 1.0, 0.0);
            }
            auto& dps = m_ages[MGO_A_DEATHS_PER_SEVERE][a];
            fix(dps, m_ages[MGO_A_CRITICAL_PER_SEVERE][a] + dps > 1.0, 0.0);
            auto& dpc = m_ages[MGO_A_DEATHS_PER_CRITICAL][a];
            fix(dpc, dpc < 0.0 || dpc > 1.0, 0.0);
        }
        return corrected;
    }
    const UncertainValue<double>& scalar_value(int k) const
    {
        return m_scalars.at(size_t(k));
    }
    // the scalar parameters' values in MGO_P_* order
    std::array<double, MGO_NUM_SCALARS> scalars() const
    {
        std::array<double, MGO_NUM_SCALARS> v{};
        for (size_t k = 0; k < v.size(); ++k) {
            v[k] = m_scalars[k];
        }
        return v;
    }
    const AgeArray& age_param(int k) const
    {
        return m_ages[size_t(k)];
    }

private:
    size_t m_G;
    std::array<UncertainValue<double>, MGO_NUM_SCALARS> m_scalars;
    std::vector<AgeArray> m_ages;
    UncertainContactMatrix<FP> m_contacts;
    mgo_mio::DynamicNPIs<FP> m_npis;
};

// Compartment values indexed by {AgeGroup, InfectionState}; flat index g*10 + state. Each compartment is
// an UncertainValue (populations.h over CustomIndexArray<UncertainValue>).
class Populations
{
public:
    struct Index {
        AgeGroup group;
        InfectionState state;
    };
    explicit Populations(size_t num_groups = 0)
        : m_G(num_groups)
        , m_v(num_groups * MGO_NUM_COMPARTMENTS, UncertainValue<double>(0.0))
    {
    }
    size_t get_flat_index(const Index& i) const
    {
        return i.group.get() * MGO_NUM_COMPARTMENTS + size_t(i.state);
    }
    UncertainValue<double>& operator[](const Index& i)
    {
        return m_v.at(get_flat_index(i));
    }
    const UncertainValue<double>& operator[](const Index& i) const
    {
        return m_v.at(get_flat_index(i));
    }
    size_t numel() const
    {
        return m_v.size();
    }
    double get_total() const
    {
        double s = 0.0;
        for (const auto& x : m_v) {
            s += double(x);
        }
        return s;
    }
    double get_group_total(AgeGroup g) const
    {
        double s = 0.0;
        for (size_t c = 0; c < MGO_NUM_COMPARTMENTS; ++c) {
            s += double(m_v[g.get() * MGO_NUM_COMPARTMENTS + c]);
        }
        return s;
    }
    // sets the compartment so that all compartments sum to total (populations.h set_difference_from_total)
    void set_difference_from_total(const Index& i, double total)
    {
        (*this)[i] = 0.0;
        (*this)[i] = total - get_total();
    }
    // sets the compartment so that its age group sums to total
    template <class Category = AgeGroup>
    void set_difference_from_group_total(const Index& i, double total)
    {
        (*this)[i] = 0.0;
        (*this)[i] = total - get_group_total(i.group);
    }
    // the compartment values in flat-index order (Populations::get_compartments)
    std::vector<double> get_compartments() const
    {
        return std::vector<double>(m_v.begin(), m_v.end());
    }

private:
    size_t m_G;
    std::vector<UncertainValue<double>> m_v;
};

template <class FP = double>
class Model
{
public:
    explicit Model(int num_agegroups)
        : populations(size_t(num_agegroups))
        , parameters(AgeGroup(size_t(num_agegroups)))
    {
    }
    size_t get_num_groups() const
    {
        return parameters.get_num_groups().get();
    }
    bool apply_constraints()
    {
        return parameters.apply_constraints();
    }
    Populations populations;
    Parameters<FP> parameters;
};

template <class FP = double>
class Simulation;

} // namespace osecir

// ---------------------------------------------------------------------------------------------- graph
template <class T>
struct Node {
    template <class... Args>
    Node(int node_id, Args&&... args)
        : id(node_id)
        , property(std::forward<Args>(args)...)
    {
    }
    int id;
    T property;
};

template <class T>
struct Edge {
    template <class... Args>
    Edge(size_t start, size_t end, Args&&... args)
        : start_node_idx(start)
        , end_node_idx(end)
        , property(std::forward<Args>(args)...)
    {
    }
    size_t start_node_idx;
    size_t end_node_idx;
    T property;
};

template <class NodePropertyT, class EdgePropertyT>
class Graph
{
public:
    using NodeProperty = NodePropertyT;
    using EdgeProperty = EdgePropertyT;

    template <class... Args>
    Node<NodePropertyT>& add_node(int id, Args&&... args)
    {
        m_nodes.emplace_back(id, std::forward<Args>(args)...);
        return m_nodes.back();
    }
    // edges stay sorted by (start, end); adding an existing edge replaces it (graph.h:231-241)
    template <class... Args>
    Edge<EdgePropertyT>& add_edge(size_t start, size_t end, Args&&... args)
    {
        if (start >= m_nodes.size() || end >= m_nodes.size()) {
            throw Error(MGO_ERR_OUT_OF_RANGE, "add_edge: node index out of range");
        }
        auto it = std::lower_bound(m_edges.begin(), m_edges.end(), std::make_pair(start, end),
                                   [](const Edge<EdgePropertyT>& e, const std::pair<size_t, size_t>& k) {
                                       return std::make_pair(e.start_node_idx, e.end_node_idx) < k;
                                   });
        if (it != m_edges.end() && it->start_node_idx == start && it->end_node_idx == end) {
            *it = Edge<EdgePropertyT>(start, end, std::forward<Args>(args)...);
            return *it;
        }
        return *m_edges.insert(it, Edge<EdgePropertyT>(start, end, std::forward<Args>(args)...));
    }
    std::vector<Node<NodePropertyT>>& nodes()
    {
        return m_nodes;
    }
    const std::vector<Node<NodePropertyT>>& nodes() const
    {
        return m_nodes;
    }
    std::vector<Edge<EdgePropertyT>>& edges()
    {
        return m_edges;
    }
    const std::vector<Edge<EdgePropertyT>>& edges() const
    {
        return m_edges;
    }

private:
    std::vector<Node<NodePropertyT>> m_nodes;
    std::vector<Edge<EdgePropertyT>> m_edges;
};

// MobilityCoefficients (metapopulation_mobility_instant.h:106-110): baseline B, minimum M and dampings of one
// coefficient vector (10 * G values); the coefficients at time t are B - D(t) (B - M).
struct MobilityDamping {
    Vector coeffs;
    int level = 0;
    int type  = 0;
    double time = 0.0;
};

class MobilityCoefficients
{
public:
    MobilityCoefficients() = default;
    template <class V>
    explicit MobilityCoefficients(const V& baseline)
        : m_baseline(detail::to_vector(baseline))
        , m_minimum(m_baseline.size(), 0.0)
    {
    }
    template <class V, class W>
    MobilityCoefficients(const V& baseline, const W& minimum)
        : m_baseline(detail::to_vector(baseline))
        , m_minimum(detail::to_vector(minimum))
    {
    }
    Vector& get_baseline()
    {
        return m_baseline;
    }
    const Vector& get_baseline() const
    {
        return m_baseline;
    }
    Vector& get_minimum()
    {
        return m_minimum;
    }
    const Vector& get_minimum() const
    {
        return m_minimum;
    }
    // a scalar value damps every entry
    template <class V>
    void add_damping(const V& coeffs, DampingLevel level, DampingType type, SimulationTime t)
    {
        MobilityDamping d;
        if constexpr (std::is_arithmetic_v<V>) {
            d.coeffs = Vector(m_baseline.size(), double(coeffs));
        }
        else {
            d.coeffs = detail::to_vector(coeffs);
        }
        d.level = level.get();
        d.type  = type.get();
        d.time  = t.get();
        m_dampings.push_back(d);
    }
    const std::vector<MobilityDamping>& get_dampings() const
    {
        return m_dampings;
    }

private:
    Vector m_baseline, m_minimum;
    std::vector<MobilityDamping> m_dampings;
};

// MobilityCoefficientGroup: coefficient vectors whose values add up; add_damping damps every one
class MobilityCoefficientGroup
{
public:
    MobilityCoefficientGroup(size_t num_matrices, size_t num_elements)
        : m_mats(num_matrices, MobilityCoefficients(Vector(num_elements, 0.0)))
    {
    }
    MobilityCoefficients& operator[](size_t k)
    {
        return m_mats.at(k);
    }
    const MobilityCoefficients& operator[](size_t k) const
    {
        return m_mats.at(k);
    }
    size_t get_num_matrices() const
    {
        return m_mats.size();
    }
    template <class V>
    void add_damping(const V& coeffs, DampingLevel level, DampingType type, SimulationTime t)
    {
        for (auto& m : m_mats) {
            m.add_damping(coeffs, level, type, t);
        }
    }

private:
    std::vector<MobilityCoefficients> m_mats;
};

// MobilityParameters (metapopulation_mobility_instant.h:113-290): the coefficient group and the edge's dynamic
// NPIs; MobilityEdge owns them
template <class FP = double>
class MobilityParameters
{
public:
    explicit MobilityParameters(const MobilityCoefficientGroup& coeffs)
        : m_coeffs(coeffs)
    {
    }
    template <class V>
        requires(!std::is_same_v<std::decay_t<V>, MobilityCoefficientGroup>)
    explicit MobilityParameters(const V& coeffs)
        : m_coeffs(1, 0)
    {
        m_coeffs[0] = MobilityCoefficients(coeffs);
    }
    const MobilityCoefficientGroup& get_coefficients() const
    {
        return m_coeffs;
    }
    MobilityCoefficientGroup& get_coefficients()
    {
        return m_coeffs;
    }
    const DynamicNPIs<FP>& get_dynamic_npis_infected() const
    {
        return m_npis;
    }
    DynamicNPIs<FP>& get_dynamic_npis_infected()
    {
        return m_npis;
    }
    void set_dynamic_npis_infected(const DynamicNPIs<FP>& v)
    {
        m_npis = v;
    }

private:
    MobilityCoefficientGroup m_coeffs;
    DynamicNPIs<FP> m_npis;
};

template <class FP = double>
class MobilityEdge
{
public:
    explicit MobilityEdge(const MobilityParameters<FP>& p)
        : m_params(p)
    {
    }
    template <class V>
        requires(!std::is_same_v<std::decay_t<V>, MobilityParameters<FP>>)
    explicit MobilityEdge(const V& coeffs)
        : m_params(coeffs)
    {
    }
    // save_indices select commuter sums tracked by the reference; not tracked here
    template <class V, class Idx>
    MobilityEdge(const V& coeffs, const Idx&)
        : m_params(coeffs)
    {
    }
    const MobilityParameters<FP>& get_parameters() const
    {
        return m_params;
    }

private:
    MobilityParameters<FP> m_params;
};

namespace osecir
{
// ---------------------------------------------------------------------------------------------- sampling
// set_params_distributions_normal (models/ode_secir/parameter_space.h:46-118): a truncated normal
// distribution around every sampled value with bounds (1 -+ 2.6 dev_rel) * value (at least min_val: 0.001,
// 0 for seasonality and compartments), a constant one where the bounds meet, and a contact damping of value
// 0.5 at the middle of [t0, tmax] on every contact matrix. As in the reference,
// dev_rel * value is passed in the quantile's place of the (lb, ub, mean, quantile) constructor.
template <class FP>
void set_params_distributions_normal(Model<FP>& model, double t0, double tmax, double dev_rel)
{
    const double big = std::numeric_limits<double>::max();
    auto set_dist    = [dev_rel, big](UncertainValue<double>& v, double min_val = 0.001) {
        const double x  = v;
        const double lb = std::min(std::max(min_val, (1 - dev_rel * 2.6) * x), 0.1 * big);
        const double ub = std::min(std::max(min_val, (1 + dev_rel * 2.6) * x), 0.5 * big);
        const double diff = std::abs(lb - ub);
        if (diff <= 1e-12 || diff <= std::max(std::abs(lb), std::abs(ub)) * DBL_MIN) {
            v.set_distribution(ParameterDistributionConstant(std::min(std::max(min_val, x), 0.3 * big)));
        }
        else {
            v.set_distribution(ParameterDistributionNormal(lb, ub, std::min(std::max(min_val, x), 0.3 * big),
                                                           std::min(std::max(min_val, dev_rel * x), big)));
        }
    };
    auto& p = model.parameters;
    set_dist(p.template get<Seasonality<FP>>(), 0.0);
    set_dist(p.template get<ICUCapacity<FP>>());
    set_dist(p.template get<TestAndTraceCapacity<FP>>());
    const size_t G = model.get_num_groups();
    for (size_t g = 0; g < G; ++g) {
        for (auto st : {InfectionState::Exposed, InfectionState::InfectedNoSymptoms,
                        InfectionState::InfectedNoSymptomsConfirmed, InfectionState::InfectedSymptoms,
                        InfectionState::InfectedSymptomsConfirmed, InfectionState::InfectedSevere,
                        InfectionState::InfectedCritical, InfectionState::Recovered}) {
            set_dist(model.populations[{AgeGroup(g), st}], 0.0);
        }
    }
    for (size_t g = 0; g < G; ++g) {
        const AgeGroup a(g);
        set_dist(p.template get<TimeExposed<FP>>()[a]);
        set_dist(p.template get<TimeInfectedNoSymptoms<FP>>()[a]);
        set_dist(p.template get<TimeInfectedSymptoms<FP>>()[a]);
        set_dist(p.template get<TimeInfectedSevere<FP>>()[a]);
        set_dist(p.template get<TimeInfectedCritical<FP>>()[a]);
        set_dist(p.template get<TransmissionProbabilityOnContact<FP>>()[a]);
        set_dist(p.template get<RelativeTransmissionNoSymptoms<FP>>()[a]);
        set_dist(p.template get<RecoveredPerInfectedNoSymptoms<FP>>()[a]);
        set_dist(p.template get<RiskOfInfectionFromSymptomatic<FP>>()[a]);
        set_dist(p.template get<MaxRiskOfInfectionFromSymptomatic<FP>>()[a]);
        set_dist(p.template get<DeathsPerCritical<FP>>()[a]);
        set_dist(p.template get<SeverePerInfectedSymptoms<FP>>()[a]);
        set_dist(p.template get<CriticalPerSevere<FP>>()[a]);
    }
    auto& cp = p.template get<ContactPatterns<FP>>();
    std::vector<size_t> mats(cp.get_cont_freq_mat().get_num_matrices());
    for (size_t k = 0; k < mats.size(); ++k) {
        mats[k] = k;
    }
    cp.get_dampings().emplace_back(UncertainValue<FP>(0.5), DampingLevel(0), DampingType(0),
                                   SimulationTime(t0 + (tmax - t0) / 2), mats, Vector(G, 1.0));
    // the reference gives the distribution to the first contact damping sampling (the new one only if
    // there was none before)
    set_dist(cp.get_dampings()[0].get_value(), 0.0);
}

// draw_sample_demographics (parameter_space.h:120-151): capacities and the initial compartments E, INS, ISy,
// ISev, ICr, R are drawn; S takes the difference so that each age group keeps its total.
template <class FP>
void draw_sample_demographics(Model<FP>& model)
{
    auto& p = model.parameters;
    p.template get<ICUCapacity<FP>>().draw_sample();
    p.template get<TestAndTraceCapacity<FP>>().draw_sample();
    for (size_t g = 0; g < model.get_num_groups(); ++g) {
        const AgeGroup a(g);
        const double total = model.populations.get_group_total(a);
        for (auto st : {InfectionState::Exposed, InfectionState::InfectedNoSymptoms, InfectionState::InfectedSymptoms,
                        InfectionState::InfectedSevere, InfectionState::InfectedCritical, InfectionState::Recovered}) {
            model.populations[{a, st}].draw_sample();
        }
        model.populations.set_difference_from_group_total({a, InfectionState::Susceptible}, total);
        model.populations.set_difference_from_group_total({a, InfectionState::Susceptible},
                                                          model.populations.get_group_total(a));
    }
}

// draw_sample_infection (parameter_space.h:153-195): seasonality and the parameters shared by all age
// groups are drawn once (group 0) and copied to every group; the age-specific ones are drawn per group.
template <class FP>
void draw_sample_infection(Model<FP>& model)
{
    auto& p = model.parameters;
    const AgeGroup a0(0);
    p.template get<Seasonality<FP>>().draw_sample();
    p.template get<TimeExposed<FP>>()[a0].draw_sample();
    p.template get<TimeInfectedNoSymptoms<FP>>()[a0].draw_sample();
    p.template get<TimeInfectedSymptoms<FP>>()[a0].draw_sample();
    p.template get<RelativeTransmissionNoSymptoms<FP>>()[a0].draw_sample();
    p.template get<RiskOfInfectionFromSymptomatic<FP>>()[a0].draw_sample();
    p.template get<MaxRiskOfInfectionFromSymptomatic<FP>>()[a0].draw_sample();
    for (size_t g = 0; g < model.get_num_groups(); ++g) {
        const AgeGroup a(g);
        p.template get<TimeExposed<FP>>()[a]            = p.template get<TimeExposed<FP>>()[a0];
        p.template get<TimeInfectedNoSymptoms<FP>>()[a] = p.template get<TimeInfectedNoSymptoms<FP>>()[a0];
        p.template get<RelativeTransmissionNoSymptoms<FP>>()[a] =
            p.template get<RelativeTransmissionNoSymptoms<FP>>()[a0];
        p.template get<RiskOfInfectionFromSymptomatic<FP>>()[a] =
            p.template get<RiskOfInfectionFromSymptomatic<FP>>()[a0];
        p.template get<MaxRiskOfInfectionFromSymptomatic<FP>>()[a] =
            p.template get<MaxRiskOfInfectionFromSymptomatic<FP>>()[a0];
        p.template get<TimeInfectedSymptoms<FP>>()[a].draw_sample();
        p.template get<TimeInfectedSevere<FP>>()[a].draw_sample();
        p.template get<TimeInfectedCritical<FP>>()[a].draw_sample();
        p.template get<TransmissionProbabilityOnContact<FP>>()[a].draw_sample();
        p.template get<RecoveredPerInfectedNoSymptoms<FP>>()[a].draw_sample();
        p.template get<DeathsPerCritical<FP>>()[a].draw_sample();
        p.template get<SeverePerInfectedSymptoms<FP>>()[a].draw_sample();
        p.template get<CriticalPerSevere<FP>>()[a].draw_sample();
    }
}

// draw_sample of a model (parameter_space.h:197-209): infection and demographic parameters, the contact
// matrices from the sampled dampings, then apply_constraints.
template <class FP>
void draw_sample(Model<FP>& model)
{
    draw_sample_infection(model);
    draw_sample_demographics(model);
    model.parameters.template get<ContactPatterns<FP>>().draw_sample();
    model.apply_constraints();
}

// draw_sample of a parameter graph (parameter_space.h:211-251): node 0's infection parameters, contact
// dampings and dynamic NPIs are drawn once and given to every node; each node keeps its own demographics,
// ICU and test-and-trace capacities and school holidays, and its contact matrices are rebuilt from the
// sampled dampings; every edge gets the sampled contact dampings as mobility dampings and node 0's dynamic
// NPIs. The graph passed in is sampled in place too (its models), as in the reference.
template <class FP>
Graph<Model<FP>, MobilityParameters<FP>> draw_sample(Graph<Model<FP>, MobilityParameters<FP>>& graph)
{
    Graph<Model<FP>, MobilityParameters<FP>> sampled;
    if (graph.nodes().empty()) {
        return sampled;
    }
    auto& shared = graph.nodes()[0].property;
    draw_sample_infection(shared);
    auto& shared_contacts = shared.parameters.template get<ContactPatterns<FP>>();
    shared_contacts.draw_sample_dampings();
    auto& shared_npis = shared.parameters.template get<DynamicNPIsInfectedSymptoms<FP>>();
    shared_npis.draw_sample();
    for (auto& node : graph.nodes()) {
        auto& m = node.property;
        draw_sample_demographics(m);
        const UncertainValue<double> icu = m.parameters.template get<ICUCapacity<FP>>();
        const UncertainValue<double> tnt = m.parameters.template get<TestAndTraceCapacity<FP>>();
        const auto holidays = m.parameters.template get<ContactPatterns<FP>>().get_school_holidays();
        if (&m != &shared) {
            m.parameters = shared.parameters;
        }
        m.parameters.template get<ICUCapacity<FP>>()                           = icu;
        m.parameters.template get<TestAndTraceCapacity<FP>>()                  = tnt;
        m.parameters.template get<ContactPatterns<FP>>().get_school_holidays() = holidays;
        m.parameters.template get<ContactPatterns<FP>>().make_matrix();
        m.apply_constraints();
        sampled.add_node(node.id, m);
    }
    for (auto& e : graph.edges()) {
        MobilityParameters<FP> ep = e.property;
        const size_t n            = ep.get_coefficients()[0].get_baseline().size();
        apply_dampings(ep.get_coefficients(), shared_contacts.get_dampings(), [n](const Vector& v) {
            return make_mobility_damping_vector(n, v);
        });
        ep.set_dynamic_npis_infected(shared_npis);
        sampled.add_edge(e.start_node_idx, e.end_node_idx, ep);
    }
    return sampled;
}

// Simulation of one model; results are daily values (see the header comment).
template <class FP>
class Simulation
{
public:
    explicit Simulation(const Model<FP>& model, FP t0 = 0.0, FP dt = 0.1)
        : m_model(model)
        , m_core(std::make_unique<ControlledStepperWrapper<FP>>())
        , m_result(model.populations.numel())
        , m_t0(t0)
        , m_dt(dt)
    {
        m_result.add_time_point(t0, model.populations.get_compartments());
    }
    Simulation(const Simulation& o)
        : m_model(o.m_model)
        , m_core(o.m_core->clone())
        , m_result(o.m_result)
        , m_t0(o.m_t0)
        , m_dt(o.m_dt)
    {
    }
    Simulation(Simulation&&)            = default;
    Simulation& operator=(Simulation&&) = default;
    Simulation& operator=(const Simulation& o)
    {
        if (this != &o) {
            *this = Simulation(o);
        }
        return *this;
    }

    void set_integrator_core(std::unique_ptr<OdeIntegratorCore<FP>>&& core)
    {
        m_core = std::move(core);
    }
    OdeIntegratorCore<FP>& get_integrator_core()
    {
        return *m_core;
    }
    const OdeIntegratorCore<FP>& get_integrator_core() const
    {
        return *m_core;
    }
    Model<FP>& get_model()
    {
        return m_model;
    }
    const Model<FP>& get_model() const
    {
        return m_model;
    }
    TimeSeries<FP>& get_result()
    {
        return m_result;
    }
    const TimeSeries<FP>& get_result() const
    {
        return m_result;
    }
    FP get_t0() const
    {
        return m_t0;
    }
    FP& get_dt()
    {
        return m_dt;
    }
    FP get_dt() const
    {
        return m_dt;
    }

    // Simulation::advance(tmax) of a stand-alone model: runs from t0 to tmax on the device in one call
    // (osecir::simulate semantics); a second call with a larger tmax is not supported.
    TimeSeries<FP>& advance(FP tmax);

private:
    Model<FP> m_model;
    std::unique_ptr<OdeIntegratorCore<FP>> m_core;
    TimeSeries<FP> m_result;
    FP m_t0, m_dt;
    bool m_advanced = false;
};
} // namespace osecir

template <class FP, class Sim>
class SimulationNode
{
public:
    template <class... Args>
        requires(!(sizeof...(Args) == 1 && (std::is_same_v<std::decay_t<Args>, SimulationNode> && ...)))
    explicit SimulationNode(Args&&... args)
        : m_sim(std::forward<Args>(args)...)
    {
    }
    Sim& get_simulation()
    {
        return m_sim;
    }
    const Sim& get_simulation() const
    {
        return m_sim;
    }
    TimeSeries<FP>& get_result()
    {
        return m_sim.get_result();
    }
    const TimeSeries<FP>& get_result() const
    {
        return m_sim.get_result();
    }
    Vector get_last_state() const
    {
        return m_sim.get_result().get_last_value();
    }

private:
    Sim m_sim;
};

template <class FP = double>
using OsecirGraph = Graph<SimulationNode<FP, osecir::Simulation<FP>>, MobilityEdge<FP>>;

namespace detail
{
// mgo_graph_desc arrays of a model graph
struct GraphArrays {
    int n = 0, G = 0;
    std::vector<double> scalars, ages, cbase, cmin, dtimes, dcoeffs, pops, ecoeffs;
    std::vector<int32_t> doffs, dlevels, dtypes, estart, eend;
    // dynamic NPIs (mgo_dynamic_npis)
    std::vector<double> npi_params, npi_thr, npi_vals, npi_weights;
    std::vector<int32_t> npi_offs, npi_doffs, npi_levels, npi_types;
    std::unique_ptr<mgo_dynamic_npis> npis; // on the heap: desc keeps pointing at it when the arrays move
    // out-edge dampings and edge NPIs per source node (add_edge), edge minimum coefficients
    std::vector<double> emin;
    std::vector<std::vector<double>> out_lists, out_npis;
    std::vector<char> out_set;
    std::vector<double> od_times, od_coeffs, onpi_params, onpi_thr, onpi_vals, onpi_weights;
    std::vector<int32_t> od_offs, od_levels, od_types, onpi_offs, onpi_doffs, onpi_levels, onpi_types;
    std::unique_ptr<mgo_dynamic_npis> out_npi_desc;
    mgo_integrator_cfg integrator{};
    double dt_node = 0.1;
    mgo_graph_desc desc{};

    template <class FP>
    void add_model(const osecir::Model<FP>& m)
    {
        const int g = int(m.get_num_groups());
        if (n == 0) {
            G = g;
            doffs.push_back(0);
        }
        else if (g != G) {
            throw Error(MGO_ERR_INVALID_VALUE, "all nodes of a graph need the same number of age groups");
        }
        if (G < 1 || G > MGO_MAX_GROUPS) {
            throw Error(MGO_ERR_UNSUPPORTED, "number of age groups outside 1.." + std::to_string(MGO_MAX_GROUPS));
        }
        const auto& p = m.parameters;
        const auto sc = p.scalars();
        scalars.insert(scalars.end(), sc.begin(), sc.end());
        for (int k = 0; k < MGO_NUM_AGE_PARAMS; ++k) {
            for (int a = 0; a < G; ++a) {
                ages.push_back(p.age_param(k)[AgeGroup(size_t(a))]);
            }
        }
        // fold the contact matrix group into one matrix: at most one matrix may carry dampings (static or
        // added by dynamic NPIs)
        const ContactMatrixGroup& cg = p.template get<osecir::ContactPatterns<FP>>();
        int damped                   = add_npis(p.template get<osecir::DynamicNPIsInfectedSymptoms<FP>>(),
                                                int(cg.get_num_matrices()));
        for (size_t k = 0; k < cg.get_num_matrices(); ++k) {
            if (!cg[k].get_dampings().empty()) {
                if (damped >= 0 && damped != int(k)) {
                    throw Error(MGO_ERR_UNSUPPORTED, "more than one damped contact matrix in a node");
                }
                damped = int(k);
            }
        }
        std::vector<double> B(size_t(G) * G, 0.0), M(size_t(G) * G, 0.0);
        for (size_t k = 0; k < cg.get_num_matrices(); ++k) {
            const Matrix& b  = cg[k].get_baseline();
            const Matrix& mn = cg[k].get_minimum();
            if (b.rows() != size_t(G) || b.cols() != size_t(G) || mn.rows() != size_t(G) || mn.cols() != size_t(G)) {
                throw Error(MGO_ERR_INVALID_VALUE, "contact matrix shape differs from the number of age groups");
            }
            for (int i = 0; i < G; ++i) {
                for (int j = 0; j < G; ++j) {
                    B[size_t(i) * G + j] += b(size_t(i), size_t(j));
                    // undamped matrices add their baseline to M so that B - M keeps only the damped part
                    M[size_t(i) * G + j] += (cg.get_num_matrices() == 1 || int(k) == damped)
                                                ? mn(size_t(i), size_t(j))
                                                : b(size_t(i), size_t(j));
                }
            }
        }
        if (cg.get_num_matrices() > 1 && damped < 0) {
            M.assign(M.size(), 0.0);
            for (size_t k = 0; k < cg.get_num_matrices(); ++k) {
                for (int i = 0; i < G; ++i) {
                    for (int j = 0; j < G; ++j) {
                        M[size_t(i) * G + j] += cg[k].get_minimum()(size_t(i), size_t(j));
                    }
                }
            }
        }
        cbase.insert(cbase.end(), B.begin(), B.end());
        cmin.insert(cmin.end(), M.begin(), M.end());
        if (damped >= 0) {
            for (const Damping& d : cg[size_t(damped)].get_dampings()) {
                if (d.coeffs.rows() != size_t(G) || d.coeffs.cols() != size_t(G)) {
                    throw Error(MGO_ERR_INVALID_VALUE, "damping shape differs from the number of age groups");
                }
                dtimes.push_back(d.time);
                dlevels.push_back(d.level);
                dtypes.push_back(d.type);
                dcoeffs.insert(dcoeffs.end(), d.coeffs.data(), d.coeffs.data() + size_t(G) * G);
            }
        }
        doffs.push_back(int32_t(dtimes.size()));
        const auto comp = m.populations.get_compartments();
        pops.insert(pops.end(), comp.begin(), comp.end());
        ++n;
    }

    // the node's DynamicNPIsInfectedSymptoms; returns the contact matrix they damp (-1: none)
    template <class FP>
    int add_npis(const DynamicNPIs<FP>& dn, int num_mats)
    {
        if (npi_offs.empty()) {
            npi_offs.push_back(0);
            npi_doffs.push_back(0);
        }
        npi_params.push_back(dn.get_duration().get());
        npi_params.push_back(dn.get_implementation_delay().get());
        npi_params.push_back(dn.get_base_value());
        npi_params.push_back(dn.get_directive_begin().get());
        npi_params.push_back(dn.get_directive_end().get());
        int target = -1;
        for (const auto& t : dn.get_thresholds()) {
            npi_thr.push_back(t.first);
            for (const auto& d : t.second) {
                const auto& mi = d.get_matrix_indices();
                if (mi.size() != 1 || int(mi[0]) >= num_mats || (target >= 0 && int(mi[0]) != target)) {
                    throw Error(MGO_ERR_UNSUPPORTED, "dynamic NPIs of a node must damp one and the same contact matrix");
                }
                if (d.get_group_weights().size() != size_t(G)) {
                    throw Error(MGO_ERR_INVALID_VALUE, "NPI group weights need one value per age group");
                }
                target = int(mi[0]);
                npi_vals.push_back(d.get_value());
                npi_levels.push_back(d.get_level().get());
                npi_types.push_back(d.get_type().get());
                npi_weights.insert(npi_weights.end(), d.get_group_weights().data(),
                                   d.get_group_weights().data() + G);
            }
            npi_doffs.push_back(int32_t(npi_vals.size()));
        }
        npi_offs.push_back(int32_t(npi_thr.size()));
        return target;
    }

    template <class FP>
    void add_core(const OdeIntegratorCore<FP>& core, double dt)
    {
        const mgo_integrator_cfg c = core.cfg();
        if (n == 1) {
            integrator = c;
            dt_node    = dt;
        }
        else if (c.kind != integrator.kind || c.abs_tol != integrator.abs_tol || c.rel_tol != integrator.rel_tol ||
                 c.dt_min != integrator.dt_min || c.dt_max != integrator.dt_max || dt != dt_node) {
            throw Error(MGO_ERR_UNSUPPORTED, "nodes of one graph need the same integrator core and initial step");
        }
    }

    void add_edge(size_t s, size_t e, const Vector& c)
    {
        add_edge(s, e, MobilityParameters<double>(c));
    }

    // the folded coefficient group (B, M), and the source node's out-edge dampings (per age group) and edge
    // NPIs, which all out-edges of a node share on the device
    template <class FP>
    void add_edge(size_t s, size_t e, const MobilityParameters<FP>& p)
    {
        const size_t nv             = size_t(G) * MGO_NUM_COMPARTMENTS;
        const MobilityCoefficientGroup& grp = p.get_coefficients();
        int damped                  = -1;
        for (size_t k = 0; k < grp.get_num_matrices(); ++k) {
            if (grp[k].get_baseline().size() != nv || grp[k].get_minimum().size() != nv) {
                throw Error(MGO_ERR_INVALID_VALUE, "mobility coefficients need 10 * G values");
            }
            if (!grp[k].get_dampings().empty()) {
                if (damped >= 0) {
                    throw Error(MGO_ERR_UNSUPPORTED, "more than one damped mobility coefficient vector in an edge");
                }
                damped = int(k);
            }
        }
        std::vector<double> B(nv, 0.0), M(nv, 0.0);
        for (size_t k = 0; k < grp.get_num_matrices(); ++k) {
            for (size_t i = 0; i < nv; ++i) {
                B[i] += grp[k].get_baseline()[i];
                M[i] += (damped < 0 || int(k) == damped) ? grp[k].get_minimum()[i] : grp[k].get_baseline()[i];
            }
        }
        // out-edge damping list of the source as per-group values
        std::vector<double> lst;
        if (damped >= 0) {
            for (const MobilityDamping& d : grp[size_t(damped)].get_dampings()) {
                lst.push_back(d.time);
                lst.push_back(d.level);
                lst.push_back(d.type);
                for (int a = 0; a < G; ++a) {
                    const double v = d.coeffs[size_t(a) * MGO_NUM_COMPARTMENTS];
                    for (int c = 1; c < MGO_NUM_COMPARTMENTS; ++c) {
                        if (d.coeffs[size_t(a) * MGO_NUM_COMPARTMENTS + size_t(c)] != v) {
                            throw Error(MGO_ERR_UNSUPPORTED, "mobility dampings must be uniform within an age group");
                        }
                    }
                    lst.push_back(v);
                }
            }
        }
        const DynamicNPIs<FP>& dn = p.get_dynamic_npis_infected();
        std::vector<double> npi_key;
        for (const auto& t : dn.get_thresholds()) {
            npi_key.push_back(t.first);
            for (const auto& d : t.second) {
                npi_key.push_back(d.get_value());
                npi_key.push_back(d.get_level().get());
                npi_key.push_back(d.get_type().get());
                for (size_t a = 0; a < d.get_group_weights().size(); ++a) {
                    npi_key.push_back(d.get_group_weights()[a]);
                }
            }
            npi_key.push_back(std::numeric_limits<double>::quiet_NaN()); // threshold separator
        }
        npi_key.push_back(dn.get_duration().get());
        npi_key.push_back(dn.get_base_value());
        if (out_lists.size() < size_t(n)) {
            out_lists.resize(size_t(n));
            out_npis.resize(size_t(n));
            out_set.resize(size_t(n), 0);
        }
        auto same = [](const std::vector<double>& a, const std::vector<double>& b) {
            if (a.size() != b.size()) {
                return false;
            }
            for (size_t i = 0; i < a.size(); ++i) {
                if (!(a[i] == b[i]) && !(std::isnan(a[i]) && std::isnan(b[i]))) {
                    return false;
                }
            }
            return true;
        };
        if (!out_set[s]) {
            out_set[s]   = 1;
            out_lists[s] = lst;
            out_npis[s]  = npi_key;
        }
        else if (!same(out_lists[s], lst) || !same(out_npis[s], npi_key)) {
            throw Error(MGO_ERR_UNSUPPORTED, "out-edges of a node differ in their dampings or dynamic NPIs; the "
                                             "device shares one list per node");
        }
        estart.push_back(int32_t(s));
        eend.push_back(int32_t(e));
        ecoeffs.insert(ecoeffs.end(), B.begin(), B.end());
        emin.insert(emin.end(), M.begin(), M.end());
    }

    // per-node out-edge damping and NPI arrays from the lists add_edge recorded (out_npis holds per threshold
    // value, then (value, level, type, G weights) per damping, then NaN; finally duration and base value)
    void finish_out_dampings()
    {
        od_offs.assign(1, 0);
        onpi_offs.assign(1, 0);
        onpi_doffs.assign(1, 0);
        od_times.clear(), od_coeffs.clear(), od_levels.clear(), od_types.clear();
        onpi_params.clear(), onpi_thr.clear(), onpi_vals.clear(), onpi_weights.clear(), onpi_levels.clear(),
            onpi_types.clear();
        const size_t es = 3 + size_t(G);
        for (int v = 0; v < n; ++v) {
            const bool set = size_t(v) < out_set.size() && out_set[size_t(v)];
            if (set) {
                const auto& l = out_lists[size_t(v)];
                for (size_t k = 0; k + es <= l.size(); k += es) {
                    od_times.push_back(l[k]);
                    od_levels.push_back(int32_t(l[k + 1]));
                    od_types.push_back(int32_t(l[k + 2]));
                    od_coeffs.insert(od_coeffs.end(), l.begin() + long(k + 3), l.begin() + long(k + es));
                }
            }
            od_offs.push_back(int32_t(od_times.size()));
            double duration = 14.0, base = 1.0;
            if (set) {
                const auto& q = out_npis[size_t(v)];
                duration      = q[q.size() - 2];
                base          = q[q.size() - 1];
                size_t i      = 0;
                while (i + 2 < q.size()) {
                    onpi_thr.push_back(q[i++]);
                    while (!std::isnan(q[i])) {
                        onpi_vals.push_back(q[i]);
                        onpi_levels.push_back(int32_t(q[i + 1]));
                        onpi_types.push_back(int32_t(q[i + 2]));
                        onpi_weights.insert(onpi_weights.end(), q.begin() + long(i + 3), q.begin() + long(i + 3 + size_t(G)));
                        i += 3 + size_t(G);
                    }
                    ++i; // separator
                    onpi_doffs.push_back(int32_t(onpi_vals.size()));
                }
            }
            onpi_params.insert(onpi_params.end(), {duration, 0.0, base, -DBL_MAX, DBL_MAX});
            onpi_offs.push_back(int32_t(onpi_thr.size()));
        }
    }

    const mgo_graph_desc& finish()
    {
        desc                  = mgo_graph_desc{};
        desc.num_nodes        = n;
        desc.num_groups       = G;
        desc.node_scalars     = scalars.data();
        desc.age_params       = ages.data();
        desc.contact_baseline = cbase.data();
        desc.contact_minimum  = cmin.data();
        desc.damping_offsets  = doffs.data();
        desc.damping_times    = dtimes.empty() ? nullptr : dtimes.data();
        desc.damping_levels   = dlevels.empty() ? nullptr : dlevels.data();
        desc.damping_types    = dtypes.empty() ? nullptr : dtypes.data();
        desc.damping_coeffs   = dcoeffs.empty() ? nullptr : dcoeffs.data();
        desc.populations      = pops.data();
        desc.num_edges        = int32_t(estart.size());
        desc.edge_start       = estart.empty() ? nullptr : estart.data();
        desc.edge_end         = eend.empty() ? nullptr : eend.data();
        desc.edge_coeffs      = ecoeffs.empty() ? nullptr : ecoeffs.data();
        desc.edge_coeff_len   = G * MGO_NUM_COMPARTMENTS;
        finish_out_dampings();
        bool any_min = false;
        for (double m : emin) {
            any_min = any_min || m != 0.0;
        }
        desc.edge_coeff_minimum = any_min ? emin.data() : nullptr;
        if (!od_times.empty()) {
            desc.out_damping_offsets = od_offs.data();
            desc.out_damping_times   = od_times.data();
            desc.out_damping_levels  = od_levels.data();
            desc.out_damping_types   = od_types.data();
            desc.out_damping_coeffs  = od_coeffs.data();
        }
        if (!onpi_thr.empty()) {
            out_npi_desc                  = std::make_unique<mgo_dynamic_npis>();
            out_npi_desc->node_offsets    = onpi_offs.data();
            out_npi_desc->node_params     = onpi_params.data();
            out_npi_desc->thresholds      = onpi_thr.data();
            out_npi_desc->damping_offsets = onpi_doffs.data();
            out_npi_desc->damping_values  = onpi_vals.empty() ? nullptr : onpi_vals.data();
            out_npi_desc->damping_levels  = onpi_levels.empty() ? nullptr : onpi_levels.data();
            out_npi_desc->damping_types   = onpi_types.empty() ? nullptr : onpi_types.data();
            out_npi_desc->damping_weights = onpi_weights.empty() ? nullptr : onpi_weights.data();
            desc.out_dynamic_npis         = out_npi_desc.get();
        }
        if (!npi_thr.empty()) {
            npis                  = std::make_unique<mgo_dynamic_npis>();
            npis->node_offsets    = npi_offs.data();
            npis->node_params     = npi_params.data();
            npis->thresholds      = npi_thr.data();
            npis->damping_offsets = npi_doffs.data();
            npis->damping_values  = npi_vals.empty() ? nullptr : npi_vals.data();
            npis->damping_levels  = npi_levels.empty() ? nullptr : npi_levels.data();
            npis->damping_types   = npi_types.empty() ? nullptr : npi_types.data();
            npis->damping_weights = npi_weights.empty() ? nullptr : npi_weights.data();
            desc.dynamic_npis     = npis.get();
        }
        return desc;
    }
};

template <class FP>
GraphArrays graph_arrays(const OsecirGraph<FP>& g)
{
    GraphArrays a;
    if (g.nodes().empty()) {
        throw Error(MGO_ERR_INVALID_VALUE, "graph without nodes");
    }
    for (const auto& node : g.nodes()) {
        const auto& sim = node.property.get_simulation();
        a.add_model(sim.get_model());
        a.add_core(sim.get_integrator_core(), double(sim.get_dt()));
    }
    for (const auto& e : g.edges()) {
        a.add_edge(e.start_node_idx, e.end_node_idx, e.property.get_parameters());
    }
    a.finish();
    return a;
}

// daily values [day][node][10G] -> each node's result series (plus the state at t when t is not a day)
template <class FP>
void fill_results(OsecirGraph<FP>& g, const std::vector<double>& daily, int ndays, int day0, FP t,
                  const std::vector<double>* state)
{
    const size_t n  = g.nodes().size();
    const size_t nv = n ? g.nodes()[0].property.get_result().get_num_elements() : 0;
    for (size_t v = 0; v < n; ++v) {
        TimeSeries<FP>& ts = g.nodes()[v].property.get_result();
        ts.clear();
        for (int d = 0; d < ndays; ++d) {
            ts.add_time_point(FP(day0 + d), daily.data() + (size_t(d) * n + v) * nv);
        }
        if (state && (ndays == 0 || std::abs(double(day0 + ndays - 1) - double(t)) > 1e-10)) {
            ts.add_time_point(t, state->data() + v * nv);
        }
    }
}
} // namespace detail

// GraphSimulation over an osecir mobility graph (GraphSimulation<Graph<SimulationNode<Simulation>,
// MobilityEdge>>); the graph's parameters and state live on the device from the first advance on.
template <class FP, class GraphT>
class GraphSimulation
{
public:
    GraphSimulation(FP t0, FP dt, GraphT&& g)
        : m_graph(std::move(g))
        , m_t(t0)
        , m_t0(t0)
        , m_dt(dt)
    {
    }
    GraphSimulation(FP t0, FP dt, const GraphT& g)
        : m_graph(g)
        , m_t(t0)
        , m_t0(t0)
        , m_dt(dt)
    {
    }
    GraphSimulation(GraphSimulation&& o) noexcept
        : m_graph(std::move(o.m_graph))
        , m_t(o.m_t)
        , m_t0(o.m_t0)
        , m_dt(o.m_dt)
        , m_horizon(o.m_horizon)
        , m_id(std::exchange(o.m_id, 0))
    {
    }
    GraphSimulation(const GraphSimulation&)            = delete;
    GraphSimulation& operator=(const GraphSimulation&) = delete;
    ~GraphSimulation()
    {
        if (m_id) {
            mgo_graph_destroy(default_context(), m_id);
        }
    }

    // Largest tmax that later advance calls will use (sizes the device's daily output); optional.
    void set_horizon(FP tmax)
    {
        m_horizon = tmax;
    }

    void advance(FP tmax)
    {
        mgo_context* ctx = default_context();
        if (!m_id) {
            m_arrays              = detail::graph_arrays(m_graph);
            mgo_sim_cfg cfg{};
            cfg.t0                = double(m_t0);
            cfg.dt_graph          = double(m_dt);
            cfg.dt_node           = m_arrays.dt_node;
            cfg.horizon           = double(std::max(tmax, m_horizon));
            cfg.integrator        = m_arrays.integrator;
            m_cfg                 = cfg;
            detail::check(mgo_graph_create(ctx, &m_arrays.desc, &cfg, &m_id), ctx);
        }
        else if (double(tmax) > m_cfg.horizon + 1e-10) {
            throw Error(MGO_ERR_UNSUPPORTED, "advance beyond the horizon of the first advance; call set_horizon");
        }
        detail::check(mgo_graph_advance(ctx, m_id, double(tmax)), ctx);
        double t = 0;
        detail::check(mgo_graph_get_time(ctx, m_id, &t), ctx);
        m_t            = FP(t);
        const size_t n = m_graph.nodes().size();
        const int nv   = m_arrays.G * MGO_NUM_COMPARTMENTS;
        const int cap  = int(std::floor(m_cfg.horizon + 1e-10) - std::ceil(m_cfg.t0 - 1e-10)) + 1;
        std::vector<double> daily(size_t(std::max(cap, 1)) * n * size_t(nv));
        int32_t nd = 0;
        detail::check(mgo_graph_read_daily(ctx, m_id, cap, daily.data(), &nd), ctx);
        std::vector<double> state(n * size_t(nv));
        detail::check(mgo_graph_read_state(ctx, m_id, state.data()), ctx);
        detail::fill_results(m_graph, daily, nd, int(std::ceil(m_cfg.t0 - 1e-10)), m_t, &state);
    }

    FP get_t() const
    {
        return m_t;
    }
    FP get_dt() const
    {
        return m_dt;
    }
    GraphT& get_graph() &
    {
        return m_graph;
    }
    const GraphT& get_graph() const&
    {
        return m_graph;
    }
    GraphT&& get_graph() &&
    {
        return std::move(m_graph);
    }

    // used by ParameterStudy to run many simulations in one device batch
    detail::GraphArrays& batch_arrays()
    {
        m_arrays = detail::graph_arrays(m_graph);
        return m_arrays;
    }
    void set_batch_result(const std::vector<double>& daily, int ndays, int day0, FP t)
    {
        m_t = t;
        detail::fill_results(m_graph, daily, ndays, day0, t, static_cast<const std::vector<double>*>(nullptr));
    }

private:
    GraphT m_graph;
    FP m_t, m_t0, m_dt;
    FP m_horizon = 0;
    mgo_graph_id m_id = 0;
    detail::GraphArrays m_arrays;
    mgo_sim_cfg m_cfg{};
};

// make_mobility_sim(t0, dt, graph): daily exchange with dt = 0.5
template <class FP>
GraphSimulation<FP, OsecirGraph<FP>> make_mobility_sim(FP t0, FP dt, OsecirGraph<FP>&& g)
{
    return GraphSimulation<FP, OsecirGraph<FP>>(t0, dt, std::move(g));
}
template <class FP>
GraphSimulation<FP, OsecirGraph<FP>> make_mobility_sim(FP t0, FP dt, const OsecirGraph<FP>& g)
{
    return GraphSimulation<FP, OsecirGraph<FP>>(t0, dt, g);
}

// make_sampled_graph_simulation (compartments/parameter_studies.h:382-398): the simulation graph of a sampled
// parameter graph, every node starting at t0 with step dt_node_sim, exchanged every dt_graph_sim
template <class FP, class Sim = osecir::Simulation<FP>>
GraphSimulation<FP, OsecirGraph<FP>>
make_sampled_graph_simulation(const Graph<osecir::Model<FP>, MobilityParameters<FP>>& sampled_graph, FP t0,
                              FP dt_node_sim, FP dt_graph_sim)
{
    static_assert(std::is_same_v<Sim, osecir::Simulation<FP>>, "the device path runs osecir::Simulation");
    OsecirGraph<FP> g;
    for (const auto& node : sampled_graph.nodes()) {
        g.add_node(node.id, node.property, t0, dt_node_sim);
    }
    for (const auto& e : sampled_graph.edges()) {
        g.add_edge(e.start_node_idx, e.end_node_idx, e.property);
    }
    return make_mobility_sim<FP>(t0, dt_graph_sim, std::move(g));
}

namespace osecir
{
template <class FP>
TimeSeries<FP>& Simulation<FP>::advance(FP tmax)
{
    if (m_advanced) {
        throw Error(MGO_ERR_UNSUPPORTED, "a stand-alone Simulation advances once from t0 on the device");
    }
    mgo_context* ctx = default_context();
    detail::GraphArrays a;
    a.add_model(m_model);
    a.add_core(*m_core, double(m_dt));
    const mgo_graph_desc& d = a.finish();
    mgo_sim_cfg cfg{};
    cfg.t0         = double(m_t0);
    cfg.dt_graph   = 0.5;
    cfg.dt_node    = double(m_dt);
    cfg.horizon    = double(tmax);
    cfg.integrator = a.integrator;
    const int cap  = int(std::floor(double(tmax) + 1e-10) - std::ceil(double(m_t0) - 1e-10)) + 1;
    const size_t nv = m_model.populations.numel();
    std::vector<double> out(size_t(std::max(cap, 1)) * nv);
    int32_t nd = 0;
    detail::check(mgo_simulate_single(ctx, &d, &cfg, double(tmax), cap, out.data(), &nd), ctx);
    m_result.clear();
    for (int k = 0; k < nd; ++k) {
        m_result.add_time_point(FP(std::ceil(double(m_t0) - 1e-10) + k), out.data() + size_t(k) * nv);
    }
    m_advanced = true;
    return m_result;
}

// osecir::simulate(t0, tmax, dt, model[, core]): daily values of one model (model.h:385-390)
template <class FP>
TimeSeries<FP> simulate(FP t0, FP tmax, FP dt, const Model<FP>& model,
                        std::unique_ptr<OdeIntegratorCore<FP>>&& integrator_core = nullptr)
{
    Simulation<FP> sim(model, t0, dt);
    if (integrator_core) {
        sim.set_integrator_core(std::move(integrator_core));
    }
    sim.advance(tmax);
    return sim.get_result();
}
} // namespace osecir

template <class FP, class Sim, class E>
std::vector<TimeSeries<FP>> interpolate_simulation_result(const Graph<SimulationNode<FP, Sim>, E>& g,
                                                          FP abs_tol = 1e-10)
{
    std::vector<TimeSeries<FP>> r;
    for (const auto& n : g.nodes()) {
        r.push_back(interpolate_simulation_result(n.property.get_result(), abs_tol));
    }
    return r;
}

// interpolate_ensemble_results (data/analyze_result.h:77-86): per run, the interpolated node results
template <class G>
auto interpolate_ensemble_results(const std::vector<G>& ensemble)
{
    using R = decltype(interpolate_simulation_result(std::declval<const G&>()));
    std::vector<R> r;
    r.reserve(ensemble.size());
    for (const auto& g : ensemble) {
        r.push_back(interpolate_simulation_result(g));
    }
    return r;
}

namespace detail
{
// [run][node] series as values[run][node][time][element]; the input must be uniform
template <class FP>
std::vector<double> ensemble_values(const std::vector<std::vector<TimeSeries<FP>>>& e, size_t& nodes, size_t& T,
                                    size_t& E)
{
    if (e.empty() || e[0].empty() || e[0][0].get_num_time_points() == 0 || e[0][0].get_num_elements() == 0) {
        throw Error(MGO_ERR_INVALID_VALUE, "ensemble results are empty");
    }
    nodes = e[0].size();
    T     = e[0][0].get_num_time_points();
    E     = e[0][0].get_num_elements();
    std::vector<double> v;
    v.reserve(e.size() * nodes * T * E);
    for (const auto& run : e) {
        if (run.size() != nodes) {
            throw Error(MGO_ERR_INVALID_VALUE, "ensemble results not uniform: runs differ in their number of nodes");
        }
        for (const auto& ts : run) {
            if (ts.get_num_time_points() != T || ts.get_num_elements() != E) {
                throw Error(MGO_ERR_INVALID_VALUE, "ensemble results not uniform: time points or elements differ");
            }
            for (size_t t = 0; t < T; ++t) {
                v.insert(v.end(), ts.value_data(t), ts.value_data(t) + E);
            }
        }
    }
    return v;
}

template <class FP>
std::vector<TimeSeries<FP>> ensemble_series(const std::vector<TimeSeries<FP>>& times_of, const double* values,
                                            size_t nodes, size_t T, size_t E)
{
    std::vector<TimeSeries<FP>> r(nodes, TimeSeries<FP>(E));
    for (size_t n = 0; n < nodes; ++n) {
        for (size_t t = 0; t < T; ++t) {
            r[n].add_time_point(times_of[n].get_time(t), values + (n * T + t) * E);
        }
    }
    return r;
}
} // namespace detail

// ensemble_mean (data/analyze_result.h:289-315): per node, time and element the mean over runs, summed in run
// order as value / num_runs, computed on the device (mgo_ensemble_mean_values); the times are the last run's
template <class FP>
std::vector<TimeSeries<FP>> ensemble_mean(const std::vector<std::vector<TimeSeries<FP>>>& ensemble_result)
{
    size_t nodes = 0, T = 0, E = 0;
    const std::vector<double> v = detail::ensemble_values(ensemble_result, nodes, T, E);
    std::vector<double> m(nodes * T * E);
    mgo_context* ctx = default_context();
    detail::check(mgo_ensemble_mean_values(ctx, v.data(), int32_t(ensemble_result.size()), int64_t(m.size()), m.data()),
                  ctx);
    return detail::ensemble_series(ensemble_result.back(), m.data(), nodes, T, E);
}

// ensemble_percentile (data/analyze_result.h:316-343): the sorted run values at index floor(num_runs * p),
// p in (0, 1), computed on the device (mgo_ensemble_percentiles_values); the times are the first run's
template <class FP>
std::vector<TimeSeries<FP>> ensemble_percentile(const std::vector<std::vector<TimeSeries<FP>>>& ensemble_result, FP p)
{
    size_t nodes = 0, T = 0, E = 0;
    const std::vector<double> v = detail::ensemble_values(ensemble_result, nodes, T, E);
    std::vector<double> q(nodes * T * E);
    const double pd  = p;
    mgo_context* ctx = default_context();
    detail::check(mgo_ensemble_percentiles_values(ctx, v.data(), int32_t(ensemble_result.size()), int64_t(q.size()),
                                                  &pd, 1, q.data()),
                  ctx);
    return detail::ensemble_series(ensemble_result.front(), q.data(), nodes, T, E);
}

// sum_nodes (data/analyze_result.h:267-287): per run one series, the nodes summed in node order on the device
// (mgo_sum_nodes_values); the times are each run's first node's
template <class FP>
std::vector<std::vector<TimeSeries<FP>>> sum_nodes(const std::vector<std::vector<TimeSeries<FP>>>& ensemble_result)
{
    size_t nodes = 0, T = 0, E = 0;
    const std::vector<double> v = detail::ensemble_values(ensemble_result, nodes, T, E); // [run][node][t][e]
    const size_t runs = ensemble_result.size();
    std::vector<double> w(v.size()); // [run][t][node][e], the layout of the device's daily outputs
    for (size_t r = 0; r < runs; ++r) {
        for (size_t n = 0; n < nodes; ++n) {
            for (size_t t = 0; t < T; ++t) {
                std::copy_n(v.data() + ((r * nodes + n) * T + t) * E, E, w.data() + ((r * T + t) * nodes + n) * E);
            }
        }
    }
    std::vector<double> s(runs * T * E);
    mgo_context* ctx = default_context();
    detail::check(mgo_sum_nodes_values(ctx, w.data(), int32_t(runs), int32_t(T), int32_t(nodes), int32_t(E), s.data()),
                  ctx);
    std::vector<std::vector<TimeSeries<FP>>> r(runs, std::vector<TimeSeries<FP>>(1, TimeSeries<FP>(E)));
    for (size_t k = 0; k < runs; ++k) {
        for (size_t t = 0; t < T; ++t) {
            r[k][0].add_time_point(ensemble_result[k][0].get_time(t), s.data() + (k * T + t) * E);
        }
    }
    return r;
}

// ---------------------------------------------------------------------------------------------- study
namespace detail
{
template <class T>
struct is_osecir_graph_sim : std::false_type {
};
template <class FP>
struct is_osecir_graph_sim<GraphSimulation<FP, OsecirGraph<FP>>> : std::true_type {
};

// every graph simulation advanced to tmax as one device batch (mgo_ensemble_run)
template <class FP>
void run_graph_batch(std::vector<GraphSimulation<FP, OsecirGraph<FP>>>& sims, FP tmax)
{
    if (sims.empty()) {
        return;
    }
    mgo_context* ctx = default_context();
    std::vector<mgo_graph_desc> descs;
    std::vector<mgo_sim_cfg> cfgs;
    std::vector<double> tm;
    std::vector<size_t> offs;
    size_t total = 0;
    for (auto& s : sims) {
        GraphArrays& a = s.batch_arrays();
        mgo_sim_cfg c{};
        c.t0         = double(s.get_t());
        c.dt_graph   = double(s.get_dt());
        c.dt_node    = a.dt_node;
        c.horizon    = double(tmax);
        c.integrator = a.integrator;
        descs.push_back(a.desc);
        cfgs.push_back(c);
        tm.push_back(double(tmax));
        offs.push_back(total);
        const int nd = int(std::floor(double(tmax) + 1e-10) - std::ceil(c.t0 - 1e-10)) + 1;
        total += size_t(std::max(nd, 0)) * size_t(a.n) * size_t(a.G) * MGO_NUM_COMPARTMENTS;
    }
    mgo_ensemble_desc ed{int32_t(sims.size()), descs.data(), cfgs.data(), tm.data()};
    std::vector<double> out(std::max<size_t>(total, 1));
    check(mgo_ensemble_run(ctx, &ed, out.data(), nullptr), ctx);
    for (size_t i = 0; i < sims.size(); ++i) {
        const int day0 = int(std::ceil(cfgs[i].t0 - 1e-10));
        const int nd   = int(std::floor(double(tmax) + 1e-10)) - day0 + 1;
        const size_t sz = size_t(std::max(nd, 0)) * size_t(descs[i].num_nodes) * size_t(descs[i].num_groups) *
                          MGO_NUM_COMPARTMENTS;
        std::vector<double> part(out.begin() + std::ptrdiff_t(offs[i]), out.begin() + std::ptrdiff_t(offs[i] + sz));
        sims[i].set_batch_result(part, std::max(nd, 0), day0, tmax);
    }
}
} // namespace detail

template <class ParameterType, typename TimeType = double, typename StepType = TimeType>
class ParameterStudy
{
public:
    using Parameters = ParameterType;
    using Time       = TimeType;
    using Step       = StepType;

    ParameterStudy(const Parameters& parameters, Time t0, Time tmax, Step dt, size_t num_runs)
        : m_parameters(parameters)
        , m_num_runs(num_runs)
        , m_t0(t0)
        , m_tmax(tmax)
        , m_dt(dt)
    {
    }

    // create_simulation(const Parameters&, Time t0, Step dt, size_t run_idx) -> Sim
    // process_simulation_result(Sim&&, size_t run_idx) -> R or void
    template <class CreateSimulationFunction, class ProcessSimulationResultFunction>
    auto run(CreateSimulationFunction&& create_simulation, ProcessSimulationResultFunction&& process_simulation_result)
    {
        return run_serial(std::forward<CreateSimulationFunction>(create_simulation),
                          std::forward<ProcessSimulationResultFunction>(process_simulation_result));
    }
    template <class CreateSimulationFunction>
    auto run(CreateSimulationFunction&& create_simulation)
    {
        return run_serial(std::forward<CreateSimulationFunction>(create_simulation));
    }

    template <class CreateSimulationFunction, class ProcessSimulationResultFunction>
    auto run_serial(CreateSimulationFunction&& create_simulation,
                    ProcessSimulationResultFunction&& process_simulation_result)
    {
        using Sim = std::decay_t<std::invoke_result_t<CreateSimulationFunction, const Parameters&, Time, Step, size_t>>;
        using R   = std::decay_t<std::invoke_result_t<ProcessSimulationResultFunction, Sim&&, size_t>>;
        std::vector<Sim> sims;
        sims.reserve(m_num_runs);
        for (size_t i = 0; i < m_num_runs; ++i) {
            sims.push_back(create_simulation(std::as_const(m_parameters), m_t0, m_dt, i));
        }
        if constexpr (detail::is_osecir_graph_sim<Sim>::value) {
            detail::run_graph_batch(sims, m_tmax);
        }
        else {
            for (auto& s : sims) {
                s.advance(m_tmax);
            }
        }
        if constexpr (std::is_void_v<R>) {
            for (size_t i = 0; i < sims.size(); ++i) {
                process_simulation_result(std::move(sims[i]), i);
            }
        }
        else {
            std::vector<R> out;
            out.reserve(sims.size());
            for (size_t i = 0; i < sims.size(); ++i) {
                out.push_back(process_simulation_result(std::move(sims[i]), i));
            }
            return out;
        }
    }
    template <class CreateSimulationFunction>
    auto run_serial(CreateSimulationFunction&& create_simulation)
    {
        using Sim = std::decay_t<std::invoke_result_t<CreateSimulationFunction, const Parameters&, Time, Step, size_t>>;
        return run_serial(std::forward<CreateSimulationFunction>(create_simulation), [](Sim&& s, size_t) {
            return std::move(s);
        });
    }

    size_t get_num_runs() const
    {
        return m_num_runs;
    }
    Time get_tmax() const
    {
        return m_tmax;
    }
    Time get_t0() const
    {
        return m_t0;
    }
    Step get_dt() const
    {
        return m_dt;
    }
    const Parameters& get_parameters() const
    {
        return m_parameters;
    }
    Parameters& get_parameters()
    {
        return m_parameters;
    }

private:
    Parameters m_parameters;
    size_t m_num_runs;
    Time m_t0, m_tmax;
    Step m_dt;
};

} // namespace mgo_mio

#ifndef MGO_MIO_NO_ALIAS
namespace mio = mgo_mio;
#endif

#endif // MGO_MIO_HPP
