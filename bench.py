#!/usr/bin/env python3
"""Benchmark of the Graph-ODE engine on one MI355X.

Headline: node-days/s of the batched ensemble (1024 parameter-study samples of a 400-node, G=6 graph,
100 days, adaptive Cash-Karp abs = rel = 1e-10) with inputs resident in HBM; one "step" is one full
ensemble run (ParameterStudy::run over all samples). Next to it: the plain single-graph GPU path on the
400-node graph, the per-kernel roofline of the node-phase kernel (HIP events on the library's stream),
the maximum relative error against the oracle on the CPU-baseline sample, and the CPU baseline (the
oracle, a restatement of the reference algorithm, timed on this host's cores on a bounded sample).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from memilio_amd import capi, synthetic  # noqa: E402
from memilio_amd.engine import Batch, Context, GraphSim  # noqa: E402

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--samples", type=int, default=1024)
    ap.add_argument("--nodes", type=int, default=400)
    ap.add_argument("--groups", type=int, default=6)
    ap.add_argument("--days", type=float, default=100.0)
    ap.add_argument("--cpu-samples", type=int, default=8)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (runs shard over GPUs with no "
              f"communication; see DESIGN.md)", file=sys.stderr)
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

    node_days = S * N * days
    with Context() as ctx:
        t = time.perf_counter()
        batch = Batch(ctx, specs, cfgs, tmaxs)
        log(f"batch resident in HBM in {time.perf_counter() - t:.1f}s")
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
            log(f"step {k}: {dt * 1e3:.1f} ms wall, {info.device_ms:.1f} ms device, steps {info.steps}, "
                f"attempts {info.attempts}, fix-up passes {info.fixup_passes}, replays {info.exact_replays}")
        total = sum(times)
        value = node_days * args.steps / total
        cpu_n = min(args.cpu_samples, S)
        gpu_sample = batch.read_daily(0, cpu_n) if not args.no_cpu else None
        batch.close()

        # plain single-graph path: the base graph alone, advance(days)
        plain = GraphSim(ctx, base, cfg, horizon=days)
        plain.advance(1.0)   # warm-up (graph capture, first launches)
        plain.close()
        plain = GraphSim(ctx, base, cfg, horizon=days)
        t0 = time.perf_counter()
        plain.advance(days)
        plain_s = time.perf_counter() - t0
        plain.close()
        plain_value = N * days / plain_s
        log(f"plain single graph: {plain_s * 1e3:.1f} ms -> {plain_value:.3e} node-days/s")

    # per-kernel timing of one profiled run (events bracket every node / edge launch)
    with Context(flags=capi.CTX_PROFILE_KERNELS) as pctx:
        pb = Batch(pctx, specs, cfgs, tmaxs)
        pinfo = pb.run()
        pb.close()
    node_ms = pinfo.node_kernel_ms / max(pinfo.node_launches, 1)
    flops_per_launch = node_flops(pinfo.steps, pinfo.attempts, G) / max(pinfo.node_launches, 1)
    achieved = flops_per_launch / (node_ms * 1e-3) / 1e12
    log(f"node kernel: {pinfo.node_launches} launches, avg {node_ms:.3f} ms, {achieved:.2f} TFLOP/s; "
        f"edge kernel avg {pinfo.edge_kernel_ms / max(pinfo.edge_launches, 1):.3f} ms")

    cpu = None
    max_rel = None
    if not args.no_cpu:
        from oracle import oracle_py
        sub = specs[:cpu_n]
        t0 = time.perf_counter()
        ref, st = oracle_py.ensemble_run(sub, cfgs[:cpu_n], tmaxs[:cpu_n], threads=args.cpu_threads)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": cpu_n * N * days / cpu_s, "unit": "node-days/s", "cores": args.cpu_threads,
               "kind": "port",
               "sample": f"{cpu_n} of the {S} samples ({cpu_n * N * days:.0f} node-days), oracle = reference "
                         f"algorithm restated in C++ (-O3, no FMA contraction), OpenMP over samples, "
                         f"{cpu_s:.1f} s"}
        max_rel = max(float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-6))) for g, r in zip(gpu_sample, ref))
        log(f"cpu baseline: {cpu['value']:.3e} node-days/s on {args.cpu_threads} threads; max rel err {max_rel:.2e}")

    res = {
        "metric": "node-days/s batched+served vs plain GPU at <=1e-6 rel err; p50/p99 latency",
        "value": value,
        "unit": "node-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, memilio_amd/synthetic.py)",
        "config": {"workload": f"ensemble of {S} samples of a {N}-node G={G} graph, {days:g} days, adaptive "
                               f"Cash-Karp 1e-10, daily exchange", "samples": S, "nodes": N, "groups": G,
                   "days": days, "edges": int(base.num_edges)},
        "plain_gpu_node_days_per_s": plain_value,
        "batched_vs_plain": value / plain_value,
        "max_rel_err_vs_oracle": max_rel,
        "roofline": {"bound": "mfma", "roof": "FP64 vector ALU (78.6 TFLOP/s = FP64 peak of MI355X)",
                     "kernel": f"node_phase<G={G},CashKarp54>", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "avg_launch_ms": node_ms, "launches": pinfo.node_launches,
                     "flops_per_launch": flops_per_launch},
        "cpu_baseline": cpu,
    }
    if cpu:
        res["vs_cpu"] = value / cpu["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
