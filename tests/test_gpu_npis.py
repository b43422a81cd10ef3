"""Dynamic NPIs (osecir DynamicNPIsInfectedSymptoms) on the device against the oracle, via the C ABI.

The oracle's restatement of osecir::Simulation::advance and implement_dynamic_npis is pinned by the
reference's known answers (tests/test_oracle_golden.py, cpp/tests/test_dynamic_npis.cpp). Bars as in
test_gpu_parity.py: fixed-step 1e-12 relative, adaptive 1e-6 relative, every compartment, every day.
"""
import numpy as np
import pytest

from memilio_amd import capi, synthetic
from memilio_amd.engine import Batch, Context, GraphSim, MgoError, Server
from tests import setups

pytestmark = pytest.mark.gpu

FLOOR = 1e-6


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), FLOOR))) if a.size else 0.0


def _graph(seed, n=30, G=3, **npi):
    return synthetic.with_npis(synthetic.graph(num_nodes=n, G=G, seed=seed, mean_degree=6.0), **npi)


# grid-aligned NPI times (integral delay and duration) and times that split node phases
ON_GRID = dict(duration=14.0, delay=3.0)
OFF_GRID = dict(duration=2.3, delay=0.7, directive_begin=2.0, directive_end=40.2)


@pytest.fixture(scope="module")
def ctx():
    c = Context()
    yield c
    c.close()


def _device_graph(ctx, spec, cfg, tmax):
    sim = GraphSim(ctx, spec, cfg, tmax)
    sim.advance(tmax)
    out = sim.daily()
    sim.close()
    return out


@pytest.mark.parametrize("npi", [ON_GRID, OFF_GRID], ids=["on_grid", "off_grid"])
def test_npi_graph_fixed_step(ctx, oracle, npi):
    spec = _graph(61, **npi)
    cfg = synthetic.fixed_cfg(0.1)
    ref, _ = oracle.graph_run(spec, cfg, 60.0)
    plain, _ = oracle.graph_run(synthetic.graph(num_nodes=30, G=3, seed=61, mean_degree=6.0), cfg, 60.0)
    assert rel_err(ref, plain) > 1e-2          # the NPIs changed the course
    got = _device_graph(ctx, spec, cfg, 60.0)
    assert rel_err(got, ref) < 1e-12


@pytest.mark.parametrize("npi", [ON_GRID, OFF_GRID], ids=["on_grid", "off_grid"])
def test_npi_graph_adaptive(ctx, oracle, npi):
    spec = _graph(62, G=4, **npi)
    cfg = synthetic.ck_cfg(1e-8)
    ref, _ = oracle.graph_run(spec, cfg, 45.0)
    got = _device_graph(ctx, spec, cfg, 45.0)
    assert rel_err(got, ref) < 1e-6


def test_npi_single_model(ctx, oracle):
    """osecir::simulate with dynamic NPIs: the advance runs in pieces of at most a day with a check each."""
    spec = synthetic.with_npis(synthetic.single_model(seed=63, G=2), thresholds=((100.0, 0.5), (400.0, 0.8)),
                               duration=5.5, delay=1.5)
    for cfg, tol in ((synthetic.fixed_cfg(0.1), 1e-12), (setups.rkf45_cfg(1e-8, 1e-8, 1e-6, 1.0), 1e-6),
                     (synthetic.ck_cfg(1e-9), 1e-6)):
        ref, _ = oracle.simulate_daily(spec, cfg, 50.0)
        got = ctx.simulate_single(spec, cfg, 50.0)[:, 0, :]
        assert rel_err(got, ref) < tol


def test_npi_batch_bitwise_and_reset(ctx, oracle):
    """A batch mixing runs with and without NPIs: each run equals itself run alone bit for bit, and a
    reset restores the damping lists (a second run repeats the first exactly)."""
    specs = [_graph(70 + k, n=20 + 3 * k, **(ON_GRID if k % 2 else OFF_GRID)) for k in range(4)]
    specs.append(synthetic.graph(num_nodes=25, G=3, seed=80, mean_degree=6.0))
    cfgs = [synthetic.ck_cfg(1e-8)] * len(specs)
    tmaxs = [40.0] * len(specs)
    b = Batch(ctx, specs, cfgs, tmaxs)
    b.run()
    first = b.read_daily()
    b.reset()
    b.run()
    second = b.read_daily()
    b.close()
    for k, (x, y) in enumerate(zip(first, second)):
        assert np.array_equal(x, y), k
    for k, spec in enumerate(specs):
        alone = _device_graph(ctx, spec, cfgs[k], tmaxs[k])
        assert np.array_equal(alone, first[k]), k
    ref, _ = oracle.graph_run(specs[1], cfgs[1], 40.0)
    assert rel_err(first[1], ref) < 1e-6


def test_npi_serving(oracle):
    """A served request with dynamic NPIs runs the per-G NPI kernels: equal to its batch run bit for bit;
    scenario updates are refused for it."""
    spec = _graph(90, **OFF_GRID)
    other = synthetic.graph(num_nodes=40, G=2, seed=91, mean_degree=5.0)
    cfg = synthetic.ck_cfg(1e-8)
    with Context() as c:
        srv = Server(c)
        t1 = srv.submit(spec, cfg, 30.0)
        t2 = srv.submit(other, cfg, 30.0)
        with pytest.raises(MgoError) as e:
            srv.update(t1, 0.5)
        assert e.value.status == capi.ERR_UNSUPPORTED
        infos = srv.wait([t1, t2])
        served = srv.take(t1)
        srv.take(t2)
        srv.stop()
        assert infos[t1].status == capi.OK
        b = Batch(c, [spec], [cfg], [30.0])
        b.run()
        batched = b.read_daily()[0]
        b.close()
    assert np.array_equal(served, batched)
    ref, _ = oracle.graph_run(spec, cfg, 30.0)
    assert rel_err(served, ref) < 1e-6


# ------------------------------------------------------------------------------- out-edge dampings, edge NPIs
def _edge_graph(seed, n=30, G=3, **kw):
    return synthetic.with_edge_dampings(synthetic.graph(num_nodes=n, G=G, seed=seed, mean_degree=6.0), **kw)


def test_edge_dampings_fixed_step(ctx, oracle):
    """Mobility-coefficient dampings (with minimum coefficients) and edge dynamic NPIs: fixed step 1e-12."""
    spec = _edge_graph(64)
    cfg = synthetic.fixed_cfg(0.1)
    ref, _ = oracle.graph_run(spec, cfg, 40.0)
    got = _device_graph(ctx, spec, cfg, 40.0)
    assert rel_err(got, ref) < 1e-12


def test_edge_dampings_adaptive_with_node_npis(ctx, oracle):
    """Edge dampings and NPIs together with node NPIs (off-grid), per-edge coefficient vectors, adaptive."""
    spec = synthetic.with_npis(_edge_graph(65, G=4, dampings=((3.5, 0, 0, 0.6), (8.0, 0, 1, 0.2))), **OFF_GRID)
    c = spec.edge_coeff_matrix() * np.linspace(0.5, 1.5, spec.edge_coeffs.shape[0])[:, None]
    spec.set_edges(spec.edge_start, spec.edge_end, c)
    spec.edge_coeff_minimum = 0.1 * c
    spec._desc = None
    cfg = synthetic.ck_cfg(1e-8)
    ref, _ = oracle.graph_run(spec, cfg, 35.0)
    got = _device_graph(ctx, spec, cfg, 35.0)
    assert rel_err(got, ref) < 1e-6


def test_edge_dampings_batch_reset(ctx):
    """Edge damping lists and NPI states are restored by a batch reset."""
    specs = [_edge_graph(66 + k, n=15 + 5 * k) for k in range(3)]
    cfgs = [synthetic.ck_cfg(1e-8)] * 3
    b = Batch(ctx, specs, cfgs, [30.0] * 3)
    b.run()
    first = b.read_daily()
    b.reset()
    b.run()
    second = b.read_daily()
    b.close()
    for x, y in zip(first, second):
        assert np.array_equal(x, y)
