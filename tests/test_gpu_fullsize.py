"""Device runs at the benchmark sizes, checked through size-independent properties, plus the closed form.

  - the 1024-sample ensemble of a 400-node, 6-group graph over 100 days (adaptive Cash-Karp 1e-10): every
    run keeps its total population every day (sum over nodes and compartments, from the device's
    sum_nodes), the device's ensemble percentiles are ordered and non-negative, and one sampled run equals
    the oracle's run of the same sample to 1e-6 relative, every compartment, every day;
  - the 4096-instance uneven batch (1..2000 nodes, 1..8 groups, 30..180 days, tolerances 1e-6..1e-10):
    sampled instances keep their totals, one small instance equals the oracle;
  - transmission probability 0 makes the osecir system linear: the device's daily values equal exp(A t) y0
    (scipy's matrix exponential), fixed step and adaptive.
"""
import numpy as np
import pytest

from memilio_amd import synthetic
from memilio_amd.engine import Batch, Context, GraphSim
from tests.test_oracle_math import closed_form_daily, compartment_scale_err, no_transmission_graph

pytestmark = pytest.mark.gpu

FLOOR = 1e-6


def rel_err(a, b, floor=FLOOR):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


@pytest.fixture(scope="module")
def ctx():
    c = Context()
    yield c
    c.close()


def _totals_conserved(sums, populations):
    """sums: [day][10G] of one run; populations: the run's initial [node][G][10]."""
    per_day = sums.sum(axis=1)
    initial = populations.sum()
    return float(np.max(np.abs(per_day - initial)) / initial)


def test_full_ensemble_properties(ctx, oracle):
    base = synthetic.graph(num_nodes=400, G=6, seed=1)
    specs = synthetic.ensemble(base, 1024, seed=2)
    cfg = synthetic.ck_cfg(1e-10)
    cfgs, tmaxs = [cfg] * len(specs), [100.0] * len(specs)
    b = Batch(ctx, specs, cfgs, tmaxs)
    try:
        b.run()
        sums = b.sum_nodes()                                   # [run][day][60]
        assert sums.shape == (1024, 101, 60)
        for k in range(len(specs)):
            assert _totals_conserved(sums[k], specs[k].populations) < 1e-11, k
        pct = b.ensemble_percentiles([1e-6, 0.05, 0.5, 0.95])  # the first is the ensemble minimum
        assert pct.min() > -1e-6
        assert np.all(pct[0] <= pct[1]) and np.all(pct[1] <= pct[2]) and np.all(pct[2] <= pct[3])
        sample = 777
        got = b.read_daily(first=sample, count=1)[0]
    finally:
        b.close()
    ref, _ = oracle.graph_run(specs[sample], cfg, 100.0)
    assert rel_err(got, ref) < 1e-6


def test_full_uneven_batch_properties(ctx, oracle):
    pool = synthetic.uneven(4096, seed=3)
    specs = [p[0] for p in pool]
    cfgs = [synthetic.ck_cfg(p[2]) for p in pool]
    tmaxs = [p[1] for p in pool]
    b = Batch(ctx, specs, cfgs, tmaxs)
    try:
        b.run()
        rng = np.random.default_rng(4)
        for k in rng.choice(len(specs), 16, replace=False):
            sums = b.sum_nodes(first=int(k), count=1)[0]
            assert _totals_conserved(sums, specs[k].populations) < 1e-11, int(k)
        small = [k for k in range(len(specs)) if 20 <= specs[k].num_nodes <= 60 and specs[k].num_edges > 0]
        k = small[0]
        got = b.read_daily(first=k, count=1)[0]
    finally:
        b.close()
    ref, _ = oracle.graph_run(specs[k], cfgs[k], tmaxs[k])
    assert rel_err(got, ref) < 1e-6


@pytest.mark.parametrize("cfg,tol", [(synthetic.fixed_cfg(0.1), 1e-8), (synthetic.ck_cfg(1e-12), 1e-9)],
                         ids=["fixed", "adaptive"])
def test_closed_form_on_device(ctx, cfg, tol):
    spec = no_transmission_graph(num_nodes=6, G=3)
    sim = GraphSim(ctx, spec, cfg, 30.0)
    sim.advance(30.0)
    got = sim.daily()
    sim.close()
    ref = closed_form_daily(spec, np.arange(got.shape[0], dtype=np.float64))
    assert rel_err(got, ref) < tol


@pytest.mark.parametrize("kind", ["fixed", "adaptive"])
def test_plain_graph_path_full_size(ctx, oracle, kind):
    """The plain single-graph path (mgo_graph_create / mgo_graph_advance, make_mobility_sim(...).advance) on
    the 400-node, 6-group graph with Poisson(50) out-degree over 100 days, every compartment of every node every
    day. Adaptive Cash-Karp 1e-10: 1e-6 relative. Fixed step 0.1: 1e-12 on the scale of each compartment's
    course (its largest value over the run) and 1e-9 relative to the current value; at this horizon two
    compilations of the reference arithmetic itself (with and without FMA contraction) differ by more than 1e-12
    relative in the compartments that have decayed by orders of magnitude (test_build_drift_bounds_the_per_day_bar).
    Advancing in two calls (47.5, then 100) gives the same values as one call."""
    spec = synthetic.graph(num_nodes=400, G=6, seed=1)
    cfg = synthetic.fixed_cfg(0.1) if kind == "fixed" else synthetic.ck_cfg(1e-10)
    sim = GraphSim(ctx, spec, cfg, horizon=100.0)
    sim.advance(100.0)
    got = sim.daily()
    sim.close()
    assert got.shape == (101, 400, 60)
    ref, st = oracle.graph_run(spec, cfg, 100.0)
    assert st.returns > 0
    if kind == "fixed":
        assert compartment_scale_err(got, ref) < 1e-12
        assert rel_err(got, ref) < 1e-9
    else:
        assert rel_err(got, ref) < 1e-6
    sim = GraphSim(ctx, spec, cfg, horizon=100.0)
    sim.advance(47.5)
    sim.advance(100.0)
    split = sim.daily()
    sim.close()
    assert np.array_equal(split, got)
