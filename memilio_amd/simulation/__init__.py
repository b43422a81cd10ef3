"""Python mirror of `memilio.simulation` for the osecir Graph-ODE path, backed by the HIP engine.

Reference bindings mirrored (pycode/memilio-simulation/memilio/simulation/bindings/...):
AgeGroup, Damping, ContactMatrix, ContactMatrixGroup, UncertainContactMatrix, MobilityParameters,
TimeSeries (get_num_time_points / get_time / get_value / get_last_value / get_last_time / as_ndarray),
set_log_level / LogLevel and seed_random_number_generator (accepted for call compatibility; the engine
logs nothing and draws no random numbers). The model, graph and simulation types live in
`memilio_amd.simulation.osecir`.

A script written against the reference switches with
    import memilio_amd.simulation as mio
    import memilio_amd.simulation.osecir as osecir
"""
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


def seed_random_number_generator(*args):
    """Accepted for call compatibility; parameter sampling stays with the caller."""
    return None


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


class UncertainValue:
    """A parameter value (mio::UncertainValue without a distribution); `.value` reads and writes it."""

    def __init__(self, getter, setter):
        self._get, self._set = getter, setter

    @property
    def value(self):
        return self._get()

    @value.setter
    def value(self, v):
        self._set(float(v))

    def __float__(self):
        return float(self._get())

    def __repr__(self):
        return f"UncertainValue({self._get()})"


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
    sampling inputs the reference keeps beside it (kept and written to files, not sampled here)."""

    def __init__(self, group):
        self.cont_freq_mat = group
        G = group[0].num_groups if len(group) else 1
        self.dampings = []
        self.school_holiday_damping = DampingSampling(0.0, 0, 0, 0.0, [], np.zeros(G))
        self.school_holidays = []


class DampingSampling:
    """mio::DampingSampling (epidemiology/damping_sampling.h:40-230; binding damping_sampling.cpp:33-79): a
    damping of value * group_weights that a dynamic NPI adds to the listed contact matrices."""

    def __init__(self, value, level, type, time, matrix_indices, group_weights):
        self.value = value if isinstance(value, UncertainValue) else _FixedValue(float(value))
        self.level = int(level)
        self.type = int(type)
        self.time = float(time)
        self.matrix_indices = [int(k) for k in matrix_indices]
        self.group_weights = np.asarray(group_weights, dtype=np.float64).ravel().copy()


class _FixedValue(UncertainValue):
    def __init__(self, v):
        box = [v]
        super().__init__(lambda: box[0], lambda x: box.__setitem__(0, float(x)))

    def __deepcopy__(self, memo):
        return _FixedValue(self.value)


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
