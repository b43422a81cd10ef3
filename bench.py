#!/usr/bin/env python3
"""Benchmark of the Graph-ODE engine on one MI355X.

Headline: node-days/s of the batched ensemble (1024 parameter-study samples of a 400-node, G=6 graph,
100 days, adaptive Cash-Karp abs = rel = 1e-10) with inputs resident in HBM; one "step" is one full
ensemble run (ParameterStudy::run over all samples). The timed steps launch every kernel directly with
HIP events around each node-phase and edge-phase launch on the library's stream, so the node kernel's
average launch time (and its achieved FP64 rate) comes from the timed region itself. Next to it:

* the same run replayed from hipGraph-captured half-day cycles (one graph launch per cycle);
* the plain single-graph GPU path on the 400-node graph (make_mobility_sim + advance);
* the serving loop: Poisson arrivals at 50 / 200 / 800 requests/s of solves drawn from the uneven-batch
  size distribution (1..2000 nodes, 1..8 age groups, 30..180 days, rel tol 1e-6/1e-8/1e-10), 10 % of
  arrivals being contact-scaling updates U[0.3, 1] on a resident request; p50/p99 latency, node-days/s;
* the maximum relative error against the oracle on the CPU-baseline sample, and the CPU baseline (the
  oracle, a restatement of the reference algorithm, timed on this host's cores on a bounded sample).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from memilio_amd import capi, synthetic  # noqa: E402
from memilio_amd.engine import Batch, Context, GraphSim, MgoError, Server  # noqa: E402

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rhs_flops(G):
    """Algorithmic FLOPs of one right-hand side (FMA = 2, div/sin/cos/pow = 1): 2G^2 + 64G + 7."""
    return 2 * G * G + 64 * G + 7


def node_flops(steps, attempts, G):
    """Node-phase FLOPs from the step/attempt counters: per step (1 + 5A) RHS + A * 540 G stage work."""
    return steps * rhs_flops(G) + attempts * (5 * rhs_flops(G) + 540 * G)


def ensemble_bench(args, ctx, specs, cfgs, tmaxs):
    """One hipGraph-replayed run (one graph launch per half-day cycle), then the timed steps with events
    around every node / edge launch."""
    S, N, days = len(specs), specs[0].num_nodes, tmaxs[0]
    node_days = S * N * days
    out = {}
    t = time.perf_counter()
    batch = Batch(ctx, specs, cfgs, tmaxs)
    log(f"batch resident in HBM in {time.perf_counter() - t:.1f}s (autotune included)")
    out["graph_value"] = None
    if not args.no_graph:
        ctx.set_flags(0)
        batch.run()               # capture + warm-up
        batch.reset()
        t0 = time.perf_counter()
        ginfo = batch.run()
        gs = time.perf_counter() - t0
        out["graph_value"] = node_days / gs
        log(f"hipGraph-replayed run: {gs * 1e3:.1f} ms wall ({ginfo.device_ms:.1f} ms device) -> "
            f"{out['graph_value']:.3e} node-days/s")
    ctx.set_flags(capi.CTX_PROFILE_KERNELS | args.base_flags)
    for w in range(args.warmup):
        batch.reset()
        info = batch.run()
        log(f"warmup {w}: {info.device_ms:.1f} ms device")
    times, infos = [], []
    for k in range(args.steps):
        batch.reset()
        t0 = time.perf_counter()
        info = batch.run()
        dt = time.perf_counter() - t0
        times.append(dt)
        infos.append(info)
        log(f"step {k}: tile {info.tile}, {dt * 1e3:.1f} ms wall, {info.device_ms:.1f} ms device, node kernel "
            f"{info.node_kernel_ms:.1f} ms / {info.node_launches} launches, edge kernels "
            f"{info.edge_kernel_ms:.1f} ms, steps {info.steps}, attempts {info.attempts}, fix-up passes "
            f"{info.fixup_passes}, replays {info.exact_replays}")
    ctx.set_flags(args.base_flags)
    out["download_value"] = None
    if not args.no_download:
        # the same run followed by the download of every run's daily values to host memory
        batch.reset()
        t0 = time.perf_counter()
        batch.run()
        daily = batch.read_daily()
        dl = time.perf_counter() - t0
        nbytes = sum(d.nbytes for d in daily)
        del daily
        out["download_value"] = node_days / dl
        log(f"run + download of {nbytes / 1e9:.1f} GB of daily values: {dl * 1e3:.1f} ms -> "
            f"{out['download_value']:.3e} node-days/s")
    total = sum(times)
    out["value"] = node_days * args.steps / total
    out["ms_per_step"] = total / args.steps * 1e3
    out["infos"] = infos
    out["tile"] = infos[-1].tile if infos else None
    out["sample"] = batch.read_daily(0, min(args.cpu_samples, S)) if not args.no_cpu else None
    batch.close()
    return out


def plain_bench(ctx, base, cfg, days):
    plain = GraphSim(ctx, base, cfg, horizon=days)
    plain.advance(1.0)   # warm-up (graph capture, first launches)
    plain.close()
    plain = GraphSim(ctx, base, cfg, horizon=days)
    t0 = time.perf_counter()
    plain.advance(days)
    plain_s = time.perf_counter() - t0
    plain.close()
    value = base.num_nodes * days / plain_s
    log(f"plain single graph: {plain_s * 1e3:.1f} ms -> {value:.3e} node-days/s")
    return value


def uneven_bench(args, ctx):
    """Uneven batch: independent graphs with log-uniform node counts in [1, 2000], G uniform in 1..8,
    Poisson(min(50, n-1)) out-degree, horizons U[30, 180] days, rel tol in {1e-6, 1e-8, 1e-10}; one batch,
    one warm-up run, then timed runs (graph-replayed half-day cycles)."""
    t = time.perf_counter()
    pool = synthetic.uneven(args.uneven, seed=3)
    specs = [p[0] for p in pool]
    cfgs = [synthetic.ck_cfg(p[2]) for p in pool]
    tmaxs = [p[1] for p in pool]
    node_days = float(sum(sp.num_nodes * tm for sp, tm in zip(specs, tmaxs)))
    log(f"uneven batch: {len(specs)} graphs, {sum(sp.num_nodes for sp in specs)} nodes, "
        f"{sum(sp.num_edges for sp in specs)} edges, {node_days:.3e} node-days, generated in "
        f"{time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    batch = Batch(ctx, specs, cfgs, tmaxs)
    log(f"uneven batch resident in HBM in {time.perf_counter() - t:.1f}s")
    batch.run()
    times = []
    for _ in range(args.uneven_steps):
        batch.reset()
        t0 = time.perf_counter()
        info = batch.run()
        times.append(time.perf_counter() - t0)
        log(f"uneven step: {times[-1] * 1e3:.1f} ms wall, {info.device_ms:.1f} ms device, "
            f"{info.half_day_cycles} cycles, fix-up passes {info.fixup_passes}, replays {info.exact_replays}")
    batch.close()
    value = node_days * len(times) / sum(times)
    log(f"uneven batch: {value:.3e} node-days/s")
    return {"workload": f"{len(specs)} independent graphs, nodes log-uniform in [1, 2000], G in 1..8, "
                        f"horizons 30..180 days, rel tol 1e-6/1e-8/1e-10", "graphs": len(specs),
            "node_days": node_days, "node_days_per_s": value, "ms_per_run": sum(times) / len(times) * 1e3}


def serving_bench(args, ctx):
    """Poisson request streams over one serving loop per rate; requests drawn from the uneven-batch
    distribution. Latency = completion - submission on the context clock."""
    t = time.perf_counter()
    pool = synthetic.uneven(args.serve_pool, seed=3)
    log(f"serving pool of {len(pool)} requests generated in {time.perf_counter() - t:.1f}s")
    results = []
    for rate in args.serve_rates:
        rng = np.random.default_rng([7, int(rate)])
        n_arr = max(20, int(rate * args.serve_seconds))
        gaps = rng.exponential(1.0 / rate, n_arr)
        is_update = rng.random(n_arr) < 0.1
        picks = rng.integers(0, len(pool), n_arr)
        scales = rng.uniform(0.3, 1.0, n_arr)
        srv = Server(ctx, flags=capi.SERVE_BATCH_KERNELS if args.serve_batch_kernels else 0)
        live, tickets, updates = [], [], 0
        meta = {}
        # a client thread takes every result as it finishes (its pinned staging is recycled at once)
        infos, submitted_all, n_solves = {}, threading.Event(), [0]

        def client():
            pending = []
            while True:
                got = srv.poll()
                for r in got:
                    infos[r.ticket] = r
                    pending.append(r.ticket)
                keep = []
                for tk in pending:
                    if tk not in srv.requests:   # submit() has not recorded it yet
                        keep.append(tk)
                        continue
                    try:
                        srv.take(tk)
                    except MgoError:
                        pass                     # a failed request: its status is in infos
                pending = keep
                if submitted_all.is_set() and len(infos) >= n_solves[0] and not pending:
                    return
                if not got:
                    time.sleep(0.0005)

        taker = threading.Thread(target=client, daemon=True)
        taker.start()
        t_start = time.perf_counter()
        t_next = t_start
        for k in range(n_arr):
            t_next += gaps[k]
            while True:
                now = time.perf_counter()
                if now >= t_next:
                    break
                time.sleep(min(t_next - now, 0.002))
            if is_update[k]:
                live = [tk for tk in live if tk not in infos]
                if live:
                    tk = live[int(rng.integers(0, len(live)))]
                    try:
                        srv.update(tk, float(scales[k]))
                        updates += 1
                    except Exception:
                        pass   # finished between the check and the update
                continue
            spec, horizon, tol = pool[picks[k]]
            tk = srv.submit(spec, synthetic.ck_cfg(tol), horizon)
            tickets.append(tk)
            live.append(tk)
            meta[tk] = spec.num_nodes * horizon
        n_solves[0] = len(tickets)
        submitted_all.set()
        taker.join(timeout=600.0)
        if taker.is_alive():
            raise TimeoutError(f"serving at {rate} req/s: results not complete in 600 s")
        t_end = time.perf_counter()
        st = srv.stats()
        srv.stop()
        bad = [tk for tk in tickets if infos[tk].status != capi.OK]
        lat = np.array([infos[tk].done_ms - infos[tk].submit_ms for tk in tickets])
        wait = np.array([infos[tk].admit_ms - infos[tk].submit_ms for tk in tickets])
        service = np.array([infos[tk].retire_ms - infos[tk].admit_ms for tk in tickets])
        copy = np.array([infos[tk].done_ms - infos[tk].retire_ms for tk in tickets])
        first = min(infos[tk].submit_ms for tk in tickets)
        last = max(infos[tk].done_ms for tk in tickets)
        nd = float(sum(meta.values()))
        res = {"rate_rps": rate, "requests": len(tickets), "updates": updates, "failed": len(bad),
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "p50_admission_wait_ms": float(np.percentile(wait, 50)),
               "p50_resident_ms": float(np.percentile(service, 50)),
               "p50_result_copy_ms": float(np.percentile(copy, 50)),
               "node_days": nd, "node_days_per_s": nd / ((last - first) * 1e-3),
               "loop": {"iterations": st["iterations"], "cycles": st["cycles"],
                        "max_resident": st["max_resident_seen"], "device_s": st["device_ms"] * 1e-3,
                        "host_s": st["host_ms"] * 1e-3,
                        "device_node_days_per_s": 0.5 * st["node_cycles"] / max(st["device_ms"] * 1e-3, 1e-9)},
               "wall_s": t_end - t_start}
        log(f"serve {rate} req/s: {len(tickets)} solves + {updates} updates, p50 {res['p50_ms']:.1f} ms, "
            f"p99 {res['p99_ms']:.1f} ms (p50 wait for admission {res['p50_admission_wait_ms']:.1f} ms, resident "
            f"{res['p50_resident_ms']:.1f} ms, result copy {res['p50_result_copy_ms']:.1f} ms), "
            f"{res['node_days_per_s']:.3e} node-days/s, failed {len(bad)}")
        lp = res["loop"]
        log(f"  loop: {lp['iterations']} iterations, {lp['cycles']} cycles, up to {lp['max_resident']} resident, "
            f"device {lp['device_s']:.2f} s ({lp['device_node_days_per_s']:.3e} node-days/s while running), "
            f"host {lp['host_s']:.2f} s")
        results.append(res)
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--samples", type=int, default=1024)
    ap.add_argument("--nodes", type=int, default=400)
    ap.add_argument("--groups", type=int, default=6)
    ap.add_argument("--days", type=float, default=100.0)
    ap.add_argument("--cpu-samples", type=int, default=64)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="OpenMP threads of the CPU baseline (capped at the host's core count)")
    ap.add_argument("--no-download", action="store_true", help="skip the run timed with its result download")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the hipGraph-replayed comparison run")
    ap.add_argument("--tile", type=int, default=0, help="runs per node-major tile (0 = autotuned)")
    ap.add_argument("--no-graph-capture", action="store_true",
                    help="launch every kernel directly everywhere (no hipGraph capture; implies --no-graph)")
    ap.add_argument("--uneven", type=int, default=4096, help="graphs in the uneven batch (0 = skip)")
    ap.add_argument("--uneven-steps", type=int, default=2)
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[50.0, 200.0, 800.0, 1600.0])
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-pool", type=int, default=384)
    ap.add_argument("--serve-batch-kernels", action="store_true",
                    help="serve with the per-(integrator, G) batch kernels (MGO_SERVE_BATCH_KERNELS)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: every run of this build uses "
              f"a single GPU (ensembles would shard over GPUs by runs, with no communication)", file=sys.stderr)
        sys.exit(2)

    S, N, G, days = args.samples, args.nodes, args.groups, args.days
    log(f"workload: ensemble {S} x {N}-node graph, G={G}, {days:g} days")
    t = time.perf_counter()
    base = synthetic.graph(num_nodes=N, G=G, seed=1)
    specs = synthetic.ensemble(base, num_samples=S, seed=2)
    cfg = synthetic.ck_cfg(1e-10)
    cfgs = [cfg] * S
    tmaxs = [days] * S
    log(f"inputs generated in {time.perf_counter() - t:.1f}s, {base.num_edges} edges")

    if args.no_graph_capture:
        args.no_graph = True
    base_flags = capi.CTX_NO_GRAPH_CAPTURE if args.no_graph_capture else 0
    args.base_flags = base_flags
    ctx = Context(flags=base_flags)   # one context (one stream) for every measurement of this process
    ctx.set_tile(args.tile)
    ens = ensemble_bench(args, ctx, specs, cfgs, tmaxs)
    infos = ens["infos"]
    node_ms_sum = sum(i.node_kernel_ms for i in infos)
    launches = sum(i.node_launches for i in infos)
    node_ms = node_ms_sum / max(launches, 1)
    flops = sum(node_flops(i.steps, i.attempts, G) for i in infos)
    flops_per_launch = flops / max(launches, 1)
    achieved = flops / (node_ms_sum * 1e-3) / 1e12
    edge_ms = sum(i.edge_kernel_ms for i in infos) / max(sum(i.edge_launches for i in infos), 1)
    log(f"node kernel: {launches} launches, avg {node_ms:.3f} ms, {achieved:.2f} TFLOP/s; edge phase avg "
        f"{edge_ms:.3f} ms per cycle")

    # whole-job roofline of the ensemble (per node-day: F_nd FLOPs, B_nd compulsory HBM bytes)
    n_el = 10 * G
    deg = base.num_edges / N
    job_flops = (flops / max(len(infos), 1) + days * (S * base.num_edges * 131 * G + S * N * (2 * G * G + 17 * G + 7)))
    job_bytes = S * N * days * (40 * n_el + (8 * n_el + 8) * deg / S)
    t_roof = max(job_flops / (FP64_PEAK_TFLOPS * 1e12), job_bytes / (HBM_PEAK_GBS * 1e9))
    job_roof = {"flops_per_step": job_flops, "hbm_bytes_per_step": job_bytes, "t_roof_ms": t_roof * 1e3,
                "frac": t_roof / (ens["ms_per_step"] * 1e-3),
                "bound": "FP64" if job_flops / (FP64_PEAK_TFLOPS * 1e12) > job_bytes / (HBM_PEAK_GBS * 1e9) else "HBM"}
    log(f"job roofline: {job_flops:.3e} FLOP, {job_bytes:.3e} B per step, t_roof {t_roof * 1e3:.1f} ms "
        f"({job_roof['bound']}), fraction {job_roof['frac']:.3f}")

    plain_value = None if args.no_plain else plain_bench(ctx, base, cfg, days)
    single_value = None
    if not args.no_plain:
        # one G=6 model over 90 days (osecir::simulate) on the device
        one = synthetic.single_model(seed=0, G=G)
        ctx.simulate_single(one, cfg, 90.0)
        t0 = time.perf_counter()
        ctx.simulate_single(one, cfg, 90.0)
        single_value = 90.0 / (time.perf_counter() - t0)
        log(f"single model on the device: {single_value:.3e} node-days/s")
    uneven = uneven_bench(args, ctx) if args.uneven > 0 else None
    serving = None if args.no_serve else serving_bench(args, ctx)
    ctx.close()

    cpu = None
    max_rel = None
    cpu_other = None
    if not args.no_cpu:
        from oracle import oracle_py
        nproc = os.cpu_count() or 1
        threads = max(1, min(args.cpu_threads, nproc))
        cpu_n = min(args.cpu_samples, S)
        sub = specs[:cpu_n]
        t0 = time.perf_counter()
        ref, st = oracle_py.ensemble_run(sub, cfgs[:cpu_n], tmaxs[:cpu_n], threads=threads, fma=True)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": cpu_n * N * days / cpu_s, "unit": "node-days/s", "cores": threads, "host_nproc": nproc,
               "kind": "port",
               "sample": f"{cpu_n} of the {S} samples ({cpu_n * N * days:.0f} node-days, extrapolated linearly to "
                         f"the ensemble), the reference algorithm restated in C++ (oracle/oracle.cpp) built -O3 "
                         f"-march=x86-64-v3 with FMA contraction, OpenMP over samples on {threads} of {nproc} "
                         f"host threads, {cpu_s:.1f} s"}
        max_rel = max(float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-6)))
                      for g, r in zip(ens["sample"], ref[:len(ens["sample"])]))
        log(f"cpu baseline: {cpu['value']:.3e} node-days/s on {threads} of {nproc} threads; max rel err of the GPU "
            f"samples vs this build {max_rel:.2e}")
        # the other configurations on the CPU: one model on one core, the single graph (serial), as the
        # reference runs them (no OpenMP inside one simulation)
        one = synthetic.single_model(seed=0, G=G)
        t0 = time.perf_counter()
        oracle_py.simulate_daily(one, cfg, 90.0)
        c0_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle_py.graph_run(base, cfg, days)
        c1_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        oracle_py.graph_run_threads(base, cfg, days, threads, fma=True)
        c1p_s = time.perf_counter() - t0
        cpu_other = {"single_model_node_days_per_s": 90.0 / c0_s, "single_graph_node_days_per_s": N * days / c1_s,
                     "single_graph_omp_node_days_per_s": N * days / c1p_s, "single_graph_omp_threads": threads,
                     "note": "oracle (-O3, no FMA contraction), one core: one G=6 model over 90 days; the "
                             f"{N}-node graph over {days:g} days; and that graph with its node phase on {threads} "
                             "OpenMP threads (x86-64-v3 FMA build, edges sequential)"}
        log(f"cpu: one model {90.0 / c0_s:.3e} node-days/s, single graph {N * days / c1_s:.3e} node-days/s (one "
            f"core), {N * days / c1p_s:.3e} node-days/s (node phase on {threads} threads)")

    # HBM bytes per node-kernel launch of this workload, measured by rocprofv3 PMC passes (FETCH_SIZE,
    # WRITE_SIZE) and scaled by the counters' reading for a copy of known size (profiles/)
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", "node_phase_ensemble1024x400_g6_traffic.json")
    if os.path.exists(tf) and (S, N, G, days) == (1024, 400, 6, 100.0):
        with open(tf) as f:
            t = json.load(f)
        traffic = t["fetch_bytes_per_launch"] + t["write_bytes_per_launch"]
        traffic_src = f"profiles/{os.path.basename(tf)} ({t['command']})"

    value = ens["value"]
    res = {
        "metric": "node-days/s batched+served vs plain GPU at <=1e-6 rel err; p50/p99 latency",
        "value": value,
        "unit": "node-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ens["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, memilio_amd/synthetic.py)",
        "config": {"workload": f"ensemble of {S} samples of a {N}-node G={G} graph, {days:g} days, adaptive "
                               f"Cash-Karp 1e-10, daily exchange", "samples": S, "nodes": N, "groups": G,
                   "days": days, "edges": int(base.num_edges), "tile": ens["tile"]},
        "graph_replay_node_days_per_s": ens["graph_value"],
        "plain_gpu_node_days_per_s": plain_value,
        "with_download_node_days_per_s": ens["download_value"],
        "single_model_gpu_node_days_per_s": single_value,
        "cpu_other": cpu_other,
        "batched_vs_plain": value / plain_value if plain_value else None,
        "max_rel_err_vs_oracle": max_rel,
        "accuracy_ok": None if max_rel is None else bool(max_rel <= 1e-6),
        "roofline": {"bound": "fp64-valu", "roof": "FP64 vector ALU (78.6 TFLOP/s = FP64 peak of MI355X; no MFMA "
                                                   "is used), latency-limited at one wavefront per SIMD",
                     "kernel": f"node_phase<G={G},CashKarp54>", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src, "avg_launch_ms": node_ms, "launches": launches,
                     "flops_per_launch": flops_per_launch},
        "job_roofline": job_roof,
        "uneven_batch": uneven,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    if serving:
        best = max(s["node_days_per_s"] for s in serving)
        res["served_vs_plain"] = best / plain_value if plain_value else None
    if cpu:
        res["vs_cpu"] = value / cpu["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
