"""Python mirror of `memilio.simulation` for the osecir Graph-ODE path, backed by the HIP engine.

Reference bindings mirrored (pycode/memilio-simulation/memilio/simulation/bindings/...):
AgeGroup, Damping, ContactMatrix, ContactMatrixGroup, UncertainContactMatrix, MobilityParameters,
TimeSeries (get_num_time_points / get_time / get_value / get_last_value / get_last_time / as_ndarray),
UncertainValue, DampingSampling, DynamicNPIs, the parameter distributions and
seed_random_number_generator (numpy's generator: samples follow the reference's distributions, not its
random values), set_log_level / LogLevel (accepted for call compatibility; nothing is logged). The model,
graph, simulation and sampling functions live in `memilio_amd.simulation.osecir`.

A script written against the reference switches with
    import memilio_amd.simulation as mio
    import memilio_amd.simulation.osecir as osecir
"""
import copy
import enum

import numpy as np


class LogLevel(enum.IntEnum):
    Trace = 0
    Debug = 1
    Info = 2
    Warning = 3
    Error = 4
    Critical = 5
    Off = 6


def set_log_level(level):
    """Accepted for call compatibility; the engine does not log."""
    return None


# random numbers of the parameter distributions (draw_sample); numpy's PCG64, not the reference's Random123
# counter streams: samples have the reference's distributions, not its values
_rng = np.random.default_rng()


def seed_random_number_generator(*seeds):
    """Seeds the generator that parameter distributions draw from (mio::thread_local_rng().seed)."""
    global _rng
    _rng = np.random.default_rng([int(x) for x in seeds] if seeds else None)


def rng():
    return _rng


class AgeGroup:
    """Index of an age group (mio::AgeGroup)."""

    def __init__(self, i):
        self.i = int(i)

    def get(self):
        return self.i

    def __index__(self):
        return self.i

    def __eq__(self, o):
        return isinstance(o, AgeGroup) and o.i == self.i

    def __hash__(self):
        return hash(("AgeGroup", self.i))

    def __repr__(self):
        return f"AgeGroup({self.i})"


class ParameterDistribution:
    """mio::ParameterDistribution (utils/parameter_distributions.h:72-130): predefined samples are returned
    first, in order; then random samples."""

    type_name = None

    def __init__(self):
        self._predefined = []

    def add_predefined_sample(self, v):
        self._predefined.append(float(v))

    def remove_predefined_samples(self):
        self._predefined = []

    def get_predefined_samples(self):
        return list(self._predefined)

    def get_sample(self):
        if self._predefined:
            return self._predefined.pop(0)
        return self.get_rand_sample(_rng)

    def _json_fields(self):
        raise NotImplementedError

    def to_json(self):
        """{"Type": ..., fields, "PredefinedSamples": [...]} (parameter_distributions.h:363-432, 505-559, 867-917)."""
        js = {"Type": self.type_name}
        js.update(self._json_fields())
        js["PredefinedSamples"] = list(self._predefined)
        return js


class ParameterDistributionNormal(ParameterDistribution):
    """mio::ParameterDistributionNormal (parameter_distributions.h:159-427), constructed as the binding allows
    (utils/parameter_distributions.cpp:40-52): (lb, ub, mean, std_dev, quantile), (lb, ub, mean, quantile)
    or (mean, std_dev) with bounds [DBL_MIN, DBL_MAX]. A sample is redrawn up to 10 times while outside the
    bounds, then set to the bound it missed (to the lower one when the tenth redraw is inside: the
    reference's loop does so); before drawing, a mean outside the bounds becomes (ub - lb) / 2 and the
    standard deviation shrinks until mean +- quantile * std fits the bounds."""

    type_name = "Normal"

    def __init__(self, *args, **kw):
        super().__init__()
        if kw:
            # keyword forms of the binding: (lb, ub, mean, std_dev, quantile), (lb, ub, mean, quantile),
            # (mean, std_dev)
            order = {5: ("lb", "ub", "mean", "std_dev", "quantile"), 4: ("lb", "ub", "mean", "quantile"),
                     2: ("mean", "std_dev")}
            n = len(args) + len(kw)
            if n not in order:
                raise TypeError("ParameterDistributionNormal takes 2, 4 or 5 arguments")
            vals = dict(zip(order[n], args))
            vals.update(kw)
            args = tuple(vals[k] for k in order[n])
        self.lower_bound, self.upper_bound = np.finfo(np.float64).tiny, np.finfo(np.float64).max
        self.quantile = 2.5758
        if len(args) == 2:
            self.mean, self.standard_dev = float(args[0]), float(args[1])
        elif len(args) == 4:
            self.lower_bound, self.upper_bound, self.mean, self.quantile = map(float, args)
            self.standard_dev = self.upper_bound      # too high, then adapted
            self._adapt_standard_dev()
        elif len(args) == 5:
            self.lower_bound, self.upper_bound, self.mean, self.standard_dev, self.quantile = map(float, args)
            self._check_quantiles()
        else:
            raise TypeError("ParameterDistributionNormal(lb, ub, mean, std_dev, quantile), (lb, ub, mean, "
                            "quantile) or (mean, std_dev)")

    def _adapt_standard_dev(self):
        changed = False
        if self.mean + self.standard_dev * self.quantile > self.upper_bound:
            self.standard_dev = (self.upper_bound - self.mean) / self.quantile
            changed = True
        if self.mean - self.standard_dev * self.quantile < self.lower_bound:
            self.standard_dev = (self.mean - self.lower_bound) / self.quantile
            changed = True
        return changed

    def _check_quantiles(self):
        changed = False
        if self.mean < self.lower_bound or self.mean > self.upper_bound:
            self.mean = 0.5 * (self.upper_bound - self.lower_bound)
            changed = True
        return self._adapt_standard_dev() or changed

    def get_rand_sample(self, gen):
        if self.upper_bound == self.lower_bound:
            return self.lower_bound
        self._check_quantiles()
        r = gen.normal(self.mean, self.standard_dev)
        i = 0
        while (r > self.upper_bound or r < self.lower_bound) and i < 10:
            r = gen.normal(self.mean, self.standard_dev)
            i += 1
            if i == 10:
                r = self.upper_bound if r > self.upper_bound else self.lower_bound
        return float(r)

    def _json_fields(self):
        return {"Mean": self.mean, "StandardDev": self.standard_dev, "LowerBound": self.lower_bound,
                "UpperBound": self.upper_bound, "Quantile": self.quantile}


class ParameterDistributionUniform(ParameterDistribution):
    """mio::ParameterDistributionUniform (parameter_distributions.h:440-560)."""

    type_name = "Uniform"

    def __init__(self, lb, ub):
        super().__init__()
        self.lower_bound, self.upper_bound = float(lb), float(ub)

    def get_rand_sample(self, gen):
        return float(gen.uniform(self.lower_bound, self.upper_bound))

    def _json_fields(self):
        return {"LowerBound": self.lower_bound, "UpperBound": self.upper_bound}


class ParameterDistributionConstant(ParameterDistribution):
    """mio::ParameterDistributionConstant (parameter_distributions.h:810-920)."""

    type_name = "Constant"

    def __init__(self, constant):
        super().__init__()
        self.constant = float(constant)

    def get_rand_sample(self, gen):
        return self.constant

    def _json_fields(self):
        return {"Constant": self.constant}


def distribution_from_json(js):
    """A distribution object from its JSON (parameter_distributions.h:925-958); a missing field raises KeyError,
    an unknown Type ValueError."""
    if not isinstance(js, dict):
        raise ValueError("a distribution must be a JSON object")
    t = js["Type"]
    if t == "Normal":
        # the five-argument constructor, as the reference's reader (it adapts mean and spread to the bounds)
        d = ParameterDistributionNormal(float(js["LowerBound"]), float(js["UpperBound"]), float(js["Mean"]),
                                        float(js["StandardDev"]), float(js["Quantile"]))
    elif t == "Uniform":
        d = ParameterDistributionUniform(js["LowerBound"], js["UpperBound"])
    elif t == "LogNormal":
        d = ParameterDistributionLogNormal(js["LogMean"], js["LogStddev"])
    elif t == "Exponential":
        d = ParameterDistributionExponential(js["Rate"])
    elif t == "Constant":
        d = ParameterDistributionConstant(js["Constant"])
    else:
        raise ValueError(f"Type of ParameterDistribution {t} not valid.")
    for v in js["PredefinedSamples"]:
        d.add_predefined_sample(v)
    return d


class ParameterDistributionLogNormal(ParameterDistribution):
    """mio::ParameterDistributionLogNormal (parameter_distributions.h:566-690): exp of a normal sample."""

    type_name = "LogNormal"

    def __init__(self, log_mean, log_stddev):
        super().__init__()
        self.log_mean, self.log_stddev = float(log_mean), float(log_stddev)

    def get_rand_sample(self, gen):
        return float(gen.lognormal(self.log_mean, self.log_stddev))

    def _json_fields(self):
        return {"LogMean": self.log_mean, "LogStddev": self.log_stddev}


class ParameterDistributionExponential(ParameterDistribution):
    """mio::ParameterDistributionExponential (parameter_distributions.h:693-808)."""

    type_name = "Exponential"

    def __init__(self, rate):
        super().__init__()
        self.rate = float(rate)

    def get_rand_sample(self, gen):
        return float(gen.exponential(1.0 / self.rate))

    def _json_fields(self):
        return {"Rate": self.rate}


def distribution_to_json(d):
    return d.to_json() if isinstance(d, ParameterDistribution) else d


class UncertainValue:
    """mio::UncertainValue: `.value` reads and writes the value; set_distribution / get_distribution /
    draw_sample as the reference (utils/uncertain_value.h:150-188; the distribution is copied). Built on a
    value (UncertainValue(0.5)) or as a view of a value held elsewhere (a parameter array, populations)."""

    def __init__(self, getter=0.0, setter=None, dist_get=None, dist_set=None):
        if not callable(getter):
            box = [float(getter), None]
            getter, setter = (lambda: box[0]), (lambda v: box.__setitem__(0, float(v)))
            dist_get, dist_set = (lambda: box[1]), (lambda d: box.__setitem__(1, d))
        self._get, self._set = getter, setter
        self._dist_get, self._dist_set = dist_get, dist_set

    @property
    def value(self):
        return self._get()

    @value.setter
    def value(self, v):
        self._set(float(v))

    def set_distribution(self, dist):
        if self._dist_set is None:
            raise TypeError("this value takes no distribution")
        self._dist_set(copy.deepcopy(dist))

    def get_distribution(self):
        return self._dist_get() if self._dist_get is not None else None

    def draw_sample(self):
        d = self.get_distribution()
        if isinstance(d, ParameterDistribution):
            self._set(d.get_sample())
        return self._get()

    def __float__(self):
        return float(self._get())

    def __repr__(self):
        return f"UncertainValue({self._get()})"

    def __deepcopy__(self, memo):
        u = UncertainValue(self._get())
        d = self.get_distribution()
        if d is not None:
            u.set_distribution(d)
        return u


class Damping:
    """mio::Damping: coefficients (matrix or scalar), time t, level, type (damping.h:73-140)."""

    def __init__(self, coeffs, t=0.0, level=0, type=0):
        self.coeffs = np.atleast_2d(np.asarray(coeffs, dtype=np.float64))
        self.time = float(t)
        self.level = int(level)
        self.type = int(type)


class ContactMatrix:
    """mio::ContactMatrix: baseline, minimum, dampings (contact_matrix.h:190-194)."""

    def __init__(self, baseline, minimum=None):
        self.baseline = baseline
        self.minimum = np.zeros_like(self._baseline) if minimum is None else minimum
        self.dampings = []

    @property
    def baseline(self):
        return self._baseline

    @baseline.setter
    def baseline(self, b):
        self._baseline = np.atleast_2d(np.asarray(b, dtype=np.float64)).copy()

    @property
    def minimum(self):
        return self._minimum

    @minimum.setter
    def minimum(self, m):
        self._minimum = np.atleast_2d(np.asarray(m, dtype=np.float64)).copy()

    def add_damping(self, damping):
        self.dampings.append(damping)

    def clear_dampings(self):
        self.dampings = []

    @property
    def num_groups(self):
        return self._baseline.shape[0]


class ContactMatrixGroup:
    """mio::ContactMatrixGroup: a list of contact matrices whose effective values add up."""

    def __init__(self, num_matrices, num_groups):
        self.matrices = [ContactMatrix(np.zeros((num_groups, num_groups))) for _ in range(num_matrices)]

    def __getitem__(self, k):
        return self.matrices[k]

    def __setitem__(self, k, m):
        self.matrices[k] = m

    def __len__(self):
        return len(self.matrices)

    @property
    def num_matrices(self):
        return len(self.matrices)

    def add_damping(self, damping):
        for m in self.matrices:
            m.add_damping(damping)

    def clear_dampings(self):
        for m in self.matrices:
            m.clear_dampings()


class UncertainContactMatrix:
    """mio::UncertainContactMatrix (epidemiology/uncertain_matrix.h:40-60): `cont_freq_mat` is the
    ContactMatrixGroup; `dampings`, `school_holiday_damping` and `school_holidays` are the parameter-study
    sampling inputs kept beside it (draw_sample, make_matrix)."""

    def __init__(self, group):
        self.cont_freq_mat = group
        G = group[0].num_groups if len(group) else 1
        self.dampings = []
        self.school_holiday_damping = DampingSampling(0.0, 0, 0, 0.0, [], np.zeros(G))
        self.school_holidays = []

    def draw_sample_dampings(self):
        for d in self.dampings:
            d.draw_sample()
        self.school_holiday_damping.draw_sample()

    def make_matrix(self, accum=False):
        """The contact matrices with the sampled dampings (uncertain_matrix.h:172-198): without accum the
        matrices' dampings are replaced; each school holiday starts the holiday damping at its first day
        and sets it to 0 at its last."""
        if not accum:
            self.cont_freq_mat.clear_dampings()
        apply_dampings(self.cont_freq_mat, self.dampings, make_contact_damping_matrix)
        for a, b in self.school_holidays:
            d = copy.deepcopy(self.school_holiday_damping)
            d.time = float(a)
            apply_dampings(self.cont_freq_mat, [d], make_contact_damping_matrix)
            d.value.value = 0.0
            d.time = float(b)
            apply_dampings(self.cont_freq_mat, [d], make_contact_damping_matrix)
        return self.cont_freq_mat

    def draw_sample(self, accum=False):
        self.draw_sample_dampings()
        return self.make_matrix(accum)


class DampingSampling:
    """mio::DampingSampling (epidemiology/damping_sampling.h:40-230; binding damping_sampling.cpp:33-79): a
    damping of value * group_weights that a dynamic NPI (or a parameter study) adds to the listed contact
    matrices; the value may carry a distribution (draw_sample)."""

    def __init__(self, value, level, type, time, matrix_indices, group_weights):
        self.value = copy.deepcopy(value) if isinstance(value, UncertainValue) else UncertainValue(float(value))
        self.level = int(level)
        self.type = int(type)
        self.time = float(time)
        self.matrix_indices = [int(k) for k in matrix_indices]
        self.group_weights = np.asarray(group_weights, dtype=np.float64).ravel().copy()

    def draw_sample(self):
        self.value.draw_sample()


def make_contact_damping_matrix(groups):
    """d_ij = 1 - sqrt((1 - g_i)(1 - g_j)) (epidemiology/damping_sampling.h:268-290)."""
    g = np.asarray(groups, dtype=np.float64).ravel()
    return 1.0 - np.sqrt(np.outer(1.0 - g, 1.0 - g))


def make_mobility_damping_vector(size, groups):
    """Each group's value on every compartment of the group (damping_sampling.h:298-306)."""
    g = np.asarray(groups, dtype=np.float64).ravel()
    per = int(size) // g.size
    return np.repeat(g, per)


def apply_dampings(group, dampings, make_matrix):
    """Every damping sampling adds make_matrix(value * weights) to each matrix it lists
    (damping_sampling.h:255-266); `group` is a ContactMatrixGroup or a MobilityCoefficientGroup."""
    for d in dampings:
        for i in d.matrix_indices:
            m = make_matrix(d.value.value * d.group_weights)
            target = group[i]
            if isinstance(target, ContactMatrix):
                target.add_damping(Damping(m, t=d.time, level=d.level, type=d.type))
            else:
                target.add_damping(MobilityDamping(m, t=d.time, level=d.level, type=d.type))


class DynamicNPIs:
    """mio::DynamicNPIs (epidemiology/dynamic_npis.h:36-242; binding dynamic_npis.h:35-87): thresholds (cases
    per base_value, kept in descending order, an equal threshold replaced) and the NPIs they trigger."""

    def __init__(self):
        self._thresholds = []
        self.duration = 14.0
        self.implementation_delay = 0.0
        self.base_value = 1.0
        self.directive_begin = -np.finfo(np.float64).max
        self.directive_end = np.finfo(np.float64).max

    def set_threshold(self, threshold, dampings):
        threshold = float(threshold)
        k = 0
        while k < len(self._thresholds) and self._thresholds[k][0] > threshold:
            k += 1
        if k < len(self._thresholds) and not threshold > self._thresholds[k][0]:
            self._thresholds[k] = (threshold, list(dampings))
        else:
            self._thresholds.insert(k, (threshold, list(dampings)))

    @property
    def threshold(self):
        return list(self._thresholds)

    def draw_sample(self):
        """Samples the value of every damping of every threshold (dynamic_npis.h:192-199)."""
        for _, ds in self._thresholds:
            for d in ds:
                d.draw_sample()


class MobilityDamping:
    """mio::VectorDamping of mobility coefficients (binding simulation.cpp:97-99): coefficients (10 * G values
    or a scalar), time t, level, type."""

    def __init__(self, coeffs, t=0.0, level=0, type=0):
        self.coeffs = np.atleast_1d(np.asarray(coeffs, dtype=np.float64)).ravel().copy()
        self.time = float(t)
        self.level = int(level)
        self.type = int(type)


class MobilityCoefficients:
    """mio::MobilityCoefficients (metapopulation_mobility_instant.h:106-110): baseline, minimum, dampings;
    the coefficients at time t are baseline - D(t) (baseline - minimum)."""

    def __init__(self, baseline, minimum=None):
        self.baseline = np.asarray(baseline, dtype=np.float64).ravel().copy()
        self.minimum = np.zeros_like(self.baseline) if minimum is None else \
            np.asarray(minimum, dtype=np.float64).ravel().copy()
        self.dampings = []

    def add_damping(self, damping):
        self.dampings.append(damping)

    def clear_dampings(self):
        self.dampings = []


class MobilityCoefficientGroup:
    """mio::MobilityCoefficientGroup: coefficient vectors whose values add up; add_damping damps every one."""

    def __init__(self, num_matrices, num_elements):
        self.matrices = [MobilityCoefficients(np.zeros(num_elements)) for _ in range(num_matrices)]

    def __getitem__(self, k):
        return self.matrices[k]

    def __setitem__(self, k, m):
        self.matrices[k] = m

    def __len__(self):
        return len(self.matrices)

    @property
    def num_matrices(self):
        return len(self.matrices)

    def add_damping(self, damping):
        for m in self.matrices:
            m.add_damping(damping)


class MobilityParameters:
    """mio::MobilityParameters (binding mobility/metapopulation_mobility_instant.cpp:30-41): `coefficients` is
    the MobilityCoefficientGroup (constructed from a vector of 10 * G values or a group);
    `dynamic_npis_infected` the edge's DynamicNPIs (MobilityParameters::get_dynamic_npis_infected)."""

    def __init__(self, coeffs):
        if isinstance(coeffs, MobilityCoefficientGroup):
            self.coefficients = coeffs
        else:
            v = np.asarray(coeffs, dtype=np.float64).ravel()
            self.coefficients = MobilityCoefficientGroup(1, v.size)
            self.coefficients[0] = MobilityCoefficients(v)
        self.dynamic_npis_infected = DynamicNPIs()


class TimeSeries:
    """mio::TimeSeries (daily points on this engine)."""

    def __init__(self, num_elements, times=None, values=None):
        self.num_elements = int(num_elements)
        self.times = [] if times is None else list(times)
        self.values = [] if values is None else [np.asarray(v, dtype=np.float64) for v in values]

    def add_time_point(self, t, values):
        self.times.append(float(t))
        self.values.append(np.asarray(values, dtype=np.float64).copy())

    def get_num_time_points(self):
        return len(self.times)

    def get_num_elements(self):
        return self.num_elements

    def get_time(self, i):
        return self.times[i]

    def get_last_time(self):
        return self.times[-1]

    def get_value(self, i):
        return self.values[i]

    def get_last_value(self):
        return self.values[-1]

    def as_ndarray(self):
        """[1 + num_elements, num_time_points]: row 0 the times, then one row per element."""
        if not self.times:
            return np.zeros((1 + self.num_elements, 0))
        return np.vstack([np.asarray(self.times)[None, :], np.stack(self.values, axis=1)])

    def __len__(self):
        return len(self.times)


def interpolate_simulation_result(ts, abs_tol=1e-10):
    """Values at the integer days ceil(t0 - abs_tol) .. floor(tN + abs_tol), linear between the series'
    points (cpp/memilio/data/analyze_result.h:186-265)."""
    out = TimeSeries(ts.get_num_elements())
    n = ts.get_num_time_points()
    if n == 0:
        return out
    t = np.asarray(ts.times)
    v = np.stack(ts.values)
    d0 = int(np.ceil(t[0] - abs_tol))
    d1 = int(np.floor(t[-1] + abs_tol))
    k = 0
    for d in range(d0, d1 + 1):
        while k + 1 < n and t[k + 1] < d - abs_tol:
            k += 1
        if abs(t[k] - d) <= abs_tol or k + 1 >= n:
            out.add_time_point(d, v[k])
        elif abs(t[k + 1] - d) <= abs_tol:
            out.add_time_point(d, v[k + 1])
        else:
            w = (d - t[k]) / (t[k + 1] - t[k])
            out.add_time_point(d, v[k] + w * (v[k + 1] - v[k]))
    return out
