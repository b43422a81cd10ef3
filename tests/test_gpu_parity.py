"""Device kernels against the oracle (and, through it, the reference's golden fixtures), via the C ABI.

Bars: fixed-step runs match to 1e-12 relative, adaptive runs to 1e-6 relative, in every compartment at
every output day. Relative error is |gpu - ref| / max(|ref|, 1e-6) (1e-6 persons floors compartments
that are exactly or nearly empty). Integer/index work (edge order, outputs of identical runs) is
checked bit-exactly.
"""
import numpy as np
import pytest

from memilio_amd import capi, synthetic
from memilio_amd.engine import Batch, Context, GraphSim
from tests import setups

pytestmark = pytest.mark.gpu

FLOOR = 1e-6


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), FLOOR))) if a.size else 0.0


@pytest.fixture(scope="module")
def ctx():
    c = Context()
    yield c
    c.close()


def interp_csv_daily(ref, groups=1):
    """Linear interpolation of a fixture's step rows at integer days, as interpolate_simulation_result."""
    t = ref[:, 0]
    days = np.arange(np.ceil(t[0] - 1e-10), np.floor(t[-1] + 1e-10) + 1)
    cols = ref[:, 1:]
    return np.stack([np.array([np.interp(d, t, cols[:, k]) for k in range(cols.shape[1])]) for d in days])


# ------------------------------------------------------------------------------- single model fixtures
@pytest.mark.parametrize("name,spec_fn,cfg_fn,csv,groups", [
    ("default_g1", lambda: setups.secihurd_single(0.25, False), lambda: setups.ck_cfg(), "ode-secihurd-compare.csv", 1),
    ("default_g3", lambda: setups.secihurd_ageres(0.25), lambda: setups.ck_cfg(), "ode-secihurd-ageres-compare.csv", 3),
    ("rkf45_g1", lambda: setups.secihurd_single(0.3, True), lambda: setups.rkf45_cfg(1e-3, 1e-1, 0.3, 1.0),
     "secihurd-compare.csv", 1),
    ("cashkarp_g3", lambda: setups.secihurd_ageres(0.25), lambda: setups.ck_cfg(1e-4, 1e-1, 0.3, 1.0, dt=0.3),
     "secihurd-compare-cashkarp.csv", 3),
])
def test_fixture_single_model(ctx, oracle, name, spec_fn, cfg_fn, csv, groups):
    """Reference fixture set-ups (damping ramp, test-and-trace smoother, seasonality) on the device."""
    spec, cfg = spec_fn(), cfg_fn()
    gpu = ctx.simulate_single(spec, cfg, 50.0)[:, 0, :]
    ref, _ = oracle.simulate_daily(spec, cfg, 50.0)
    assert gpu.shape == ref.shape
    assert rel_err(gpu, ref) < 1e-6, name
    # directly against the golden CSV interpolated at integer days (summed over age groups)
    g = gpu.reshape(gpu.shape[0], groups, 10).sum(axis=1)
    csv_daily = interp_csv_daily(setups.load_csv(csv))
    assert rel_err(g[:, :csv_daily.shape[1]], csv_daily) < 1e-6, name


def test_single_model_adaptive(ctx, oracle):
    """Single osecir model, G=6, 90 days, Cash-Karp abs = rel = 1e-10: <= 1e-6 relative every day."""
    spec = synthetic.single_model(seed=0, G=6)
    cfg = synthetic.ck_cfg(1e-10)
    gpu = ctx.simulate_single(spec, cfg, 90.0)[:, 0, :]
    ref, _ = oracle.simulate_daily(spec, cfg, 90.0)
    assert gpu.shape == (91, 60)
    assert rel_err(gpu, ref) < 1e-6


def test_single_model_fixed_dt(ctx, oracle):
    """Fixed-step Cash-Karp, dt = 0.1: <= 1e-12 relative every day."""
    spec = synthetic.single_model(seed=5, G=6)
    cfg = synthetic.fixed_cfg(0.1)
    gpu = ctx.simulate_single(spec, cfg, 90.0)[:, 0, :]
    ref, _ = oracle.simulate_daily(spec, cfg, 90.0)
    assert rel_err(gpu, ref) < 1e-12


@pytest.mark.parametrize("G", [1, 2, 3, 4, 5, 7, 8])
def test_single_model_groups(ctx, oracle, G):
    """Every compiled age-group count against the oracle (fixed step: 1e-12)."""
    spec = synthetic.single_model(seed=10 + G, G=G)
    cfg = synthetic.fixed_cfg(0.1)
    gpu = ctx.simulate_single(spec, cfg, 20.0)[:, 0, :]
    ref, _ = oracle.simulate_daily(spec, cfg, 20.0)
    assert rel_err(gpu, ref) < 1e-12


# ------------------------------------------------------------------------------- graphs
def small_graph(n=24, G=6, seed=7, degree=6.0):
    return synthetic.graph(num_nodes=n, G=G, seed=seed, mean_degree=degree)


def test_graph_fixed_dt(ctx, oracle):
    """Graph with daily exchange and return, fixed dt 0.1, 30 days: <= 1e-12 relative."""
    spec = small_graph()
    cfg = synthetic.fixed_cfg(0.1)
    sim = GraphSim(ctx, spec, cfg, horizon=30.0)
    sim.advance(30.0)
    gpu = sim.daily()
    sim.close()
    ref, st = oracle.graph_run(spec, cfg, 30.0)
    assert gpu.shape == ref.shape == (31, spec.num_nodes, 60)
    assert st.returns > 0
    assert rel_err(gpu, ref) < 1e-12


def test_graph_adaptive(ctx, oracle):
    """Graph, adaptive Cash-Karp 1e-10, 30 days: <= 1e-6 relative."""
    spec = small_graph(seed=8)
    cfg = synthetic.ck_cfg(1e-10)
    sim = GraphSim(ctx, spec, cfg, horizon=30.0)
    sim.advance(30.0)
    gpu = sim.daily()
    ref, _ = oracle.graph_run(spec, cfg, 30.0)
    assert rel_err(gpu, ref) < 1e-6
    sim.close()


def test_graph_advance_in_pieces(ctx, oracle):
    """advance(t) called repeatedly (dt carried, pending travelers kept) equals one advance."""
    spec = small_graph(n=10, seed=9)
    cfg = synthetic.ck_cfg(1e-10)
    sim = GraphSim(ctx, spec, cfg, horizon=12.0)
    for t in (3.0, 3.5, 7.0, 12.0):
        sim.advance(t)
        assert sim.get_t() == t
    gpu = sim.daily()
    sim.close()
    ref, _ = oracle.graph_run(spec, cfg, 12.0)
    assert rel_err(gpu, ref) < 1e-6


def test_graph_all_features(ctx, oracle):
    """Test-and-trace and ICU smoothers active, dampings with levels/types, commuter detection,
    non-uniform edge coefficients, asymmetric contacts: fixed dt, 1e-12."""
    spec = small_graph(n=16, G=3, seed=11, degree=5.0)
    rng = np.random.default_rng(0)
    spec.node_scalars[:, capi.P_TEST_AND_TRACE_CAPACITY] = rng.uniform(5, 50, spec.num_nodes)
    spec.node_scalars[:, capi.P_ICU_CAPACITY] = rng.uniform(1, 20, spec.num_nodes)
    spec.node_scalars[:, capi.P_COMMUTER_NONDETECTION] = 0.6
    spec.node_scalars[:, capi.P_START_COMMUTER_DETECTION] = 2.0
    spec.node_scalars[:, capi.P_END_COMMUTER_DETECTION] = 9.0
    spec.contact_minimum = 0.2 * spec.contact_baseline
    spec.set_dampings([[(4.0, 0, 0, 0.5), (6.5, 1, 0, rng.uniform(0, 0.4, (3, 3))), (9.0, 0, 0, 0.2),
                        (9.0, 0, 1, 0.1)] if k % 2 == 0 else [] for k in range(spec.num_nodes)])
    ec = spec.edge_coeff_matrix() * rng.uniform(0.5, 1.5, (spec.num_edges, 30))
    spec.set_edges(spec.edge_start, spec.edge_end, ec)
    # a node with initial infections in every compartment so the smoothers switch on
    spec.populations[:, :, 2:9] = rng.uniform(0, 50, (spec.num_nodes, 3, 7))
    cfg = synthetic.fixed_cfg(0.1)
    sim = GraphSim(ctx, spec, cfg, horizon=14.0)
    sim.advance(14.0)
    gpu = sim.daily()
    sim.close()
    ref, _ = oracle.graph_run(spec, cfg, 14.0)
    assert rel_err(gpu, ref) < 1e-12


def test_ensemble_mixed_batch(ctx, oracle):
    """Several graphs of different G, sizes and tolerances in one batch; each result equals the oracle
    and is bitwise identical to the same run alone."""
    runs = synthetic.uneven(num_instances=6, seed=21, max_nodes=12)
    specs = [r[0] for r in runs]
    tm = [min(r[1], 35.0) for r in runs]
    cfgs = [synthetic.ck_cfg(r[2]) for r in runs]
    out, info = ctx.ensemble_run(specs, cfgs, tm)
    for k, (sp, c, t) in enumerate(zip(specs, cfgs, tm)):
        ref, _ = oracle.graph_run(sp, c, t)
        assert rel_err(out[k], ref) < 1e-6, k
        alone, _ = ctx.ensemble_run([sp], [c], [t])
        assert np.array_equal(alone[0], out[k]), k
    assert info.node_days == pytest.approx(sum(s.num_nodes * t for s, t in zip(specs, tm)))


def test_shared_structure_ensemble(ctx, oracle):
    """Parameter-study samples sharing one edge structure (one device copy), batch API."""
    base = small_graph(n=12, seed=13, degree=4.0)
    specs = synthetic.ensemble(base, num_samples=5, seed=4)
    cfgs = [synthetic.fixed_cfg(0.1)] * 5
    b = Batch(ctx, specs, cfgs, [10.0] * 5)
    b.run()
    out = b.read_daily()
    b.close()
    for k in range(5):
        ref, _ = oracle.graph_run(specs[k], cfgs[k], 10.0)
        assert rel_err(out[k], ref) < 1e-12


@pytest.mark.parametrize("kind", ["fixed", "adaptive"])
def test_graph_negative_fixups(ctx, oracle, kind):
    """Return phases with the negative-value fix-up (parallel passes until the fix-up records settle):
    fixed dt to 1e-12, adaptive to 1e-6, and the fix-ups did occur."""
    spec = setups.fixup_graph()
    cfg = synthetic.fixed_cfg(0.1) if kind == "fixed" else synthetic.ck_cfg(1e-10)
    out, info = ctx.ensemble_run([spec], [cfg], [20.0])
    ref, st = oracle.graph_run(spec, cfg, 20.0)
    assert st.negative_fixups > 0
    assert info.fixup_passes > 0
    assert rel_err(out[0], ref) < (1e-12 if kind == "fixed" else 1e-6)


def test_graph_negative_fixups_sequential(oracle):
    """The sequential replay (fallback of the parallel passes) gives the same result."""
    spec = setups.fixup_graph()
    cfg = synthetic.fixed_cfg(0.1)
    with Context(flags=capi.CTX_SEQUENTIAL_FIXUP) as c2:
        out, info = c2.ensemble_run([spec], [cfg], [20.0])
    ref, st = oracle.graph_run(spec, cfg, 20.0)
    assert info.exact_replays > 0
    assert rel_err(out[0], ref) < 1e-12


@pytest.mark.parametrize("case", ["single", "tie", "groups"])
def test_return_fixup_known_answers(ctx, oracle, case):
    """The return fix-up on the hand-derived set-ups of setups.fixup_known_answer (one edge; one underflow,
    a tie for the largest compartment, two underflows in one of two age groups), through the batched path,
    the plain graph path and the sequential replay: each equals the oracle and the hand-derived day-1 states
    to 1e-12 (Euler nodes, fixed step)."""
    spec, cfg, tmax, expected = setups.fixup_known_answer(case)
    ref, st = oracle.graph_run(spec, cfg, tmax)
    assert st.negative_fixups == (2 if case == "groups" else 1)
    want = expected.reshape(2, spec.G * 10)
    out, info = ctx.ensemble_run([spec], [cfg], [tmax])
    assert info.fixup_passes > 0
    sim = GraphSim(ctx, spec, cfg, horizon=tmax)
    sim.advance(tmax)
    plain = sim.daily()
    sim.close()
    with Context(flags=capi.CTX_SEQUENTIAL_FIXUP) as c2:
        seq, sinfo = c2.ensemble_run([spec], [cfg], [tmax])
    assert sinfo.exact_replays > 0
    for got in (out[0], plain, seq[0]):
        assert rel_err(got, ref) < 1e-12
        assert rel_err(got[1], want) < 1e-12


def test_deterministic(ctx):
    """Two identical runs are bitwise identical (no atomics in the exchange)."""
    spec = small_graph(n=20, seed=17)
    cfg = synthetic.ck_cfg(1e-10)
    a, _ = ctx.ensemble_run([spec], [cfg], [15.0])
    b, _ = ctx.ensemble_run([spec], [cfg], [15.0])
    assert np.array_equal(a[0], b[0])


# ------------------------------------------------------------------------------- node-major tiling
def _shared_ensemble(n_samples=40, n=30, G=3, seed=31):
    base = synthetic.graph(num_nodes=n, G=G, seed=seed, mean_degree=6.0)
    return synthetic.ensemble(base, num_samples=n_samples, seed=seed + 1)


@pytest.mark.parametrize("tile", [4, 16, 21, 40])
def test_tiling_is_bitwise_neutral(oracle, tile):
    """Runs sharing a graph interleaved node-major in tiles give bitwise the results of the run-major
    layout, with fix-ups, for every tile size (also a partial last tile; 21 = 64 / G, one graph node per
    wavefront)."""
    specs = _shared_ensemble()
    cfgs = [synthetic.ck_cfg(1e-8)] * len(specs)
    tmaxs = [12.0] * len(specs)
    with Context() as c1:
        c1.set_tile(1)
        ref, i1 = c1.ensemble_run(specs, cfgs, tmaxs)
    with Context() as c2:
        c2.set_tile(tile)
        out, i2 = c2.ensemble_run(specs, cfgs, tmaxs)
    assert i1.tile == 1 and i2.tile == tile
    for a, b in zip(ref, out):
        assert np.array_equal(a, b)
    r0, _ = oracle.graph_run(specs[0], cfgs[0], 12.0)
    assert rel_err(out[0], r0) < 1e-6


def test_tiling_mixed_batch_and_autotune(oracle):
    """Autotuned tile on a batch mixing a shared-graph group with independent graphs and a fixed-step
    run: every run equals its run-major result."""
    specs = _shared_ensemble(24, 20, 4, seed=41)
    cfgs = [synthetic.ck_cfg(1e-8)] * len(specs)
    tmaxs = [9.0] * len(specs)
    extra = [synthetic.graph(num_nodes=17, G=4, seed=50), synthetic.graph(num_nodes=25, G=2, seed=51)]
    specs += extra
    cfgs += [synthetic.ck_cfg(1e-6), synthetic.fixed_cfg(0.1)]
    tmaxs += [11.0, 7.0]
    with Context() as c1:
        c1.set_tile(1)
        ref, _ = c1.ensemble_run(specs, cfgs, tmaxs)
    with Context() as c2:
        out, info = c2.ensemble_run(specs, cfgs, tmaxs)
    assert info.tile in (1, 2, 4, 8, 16, 32, 64)   # the autotuner's candidates for G = 4 (64 / G = 16)
    for a, b in zip(ref, out):
        assert np.array_equal(a, b)
    r, _ = oracle.graph_run(extra[1], cfgs[-1], 7.0)
    assert rel_err(out[-1], r) < 1e-12


def test_ensemble_statistics_on_device(oracle):
    """ensemble_mean, ensemble_percentile and sum_nodes computed on the device from the resident daily
    outputs equal the oracle's restatement on the same outputs bitwise (37 runs: not a power of two)."""
    specs = _shared_ensemble(37, 15, 3, seed=61)
    cfgs = [synthetic.ck_cfg(1e-8)] * len(specs)
    tmaxs = [10.0] * len(specs)
    ps = [0.05, 0.25, 0.5, 0.75, 0.95]
    with Context() as c:
        b = Batch(c, specs, cfgs, tmaxs)
        b.run()
        daily = b.read_daily()
        mean = b.ensemble_mean()
        pct = b.ensemble_percentiles(ps)
        sums = b.sum_nodes()
        part = b.ensemble_mean(first=5, count=11)
        b.close()
    assert np.array_equal(mean, oracle.ensemble_mean(daily))
    for k, p in enumerate(ps):
        assert np.array_equal(pct[k], oracle.ensemble_percentile(daily, p)), p
    assert np.array_equal(sums, oracle.sum_nodes(daily))
    assert np.array_equal(part, oracle.ensemble_mean(daily[5:16]))
