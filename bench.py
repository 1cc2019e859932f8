"""Benchmark: statistical model checking throughput on one MI355X.

Workload: Fischer mutual exclusion with 32 processes,
query Pr[<=300](<> P0.cs), 1,000,000 runs per step, FP64 clocks and time.
One step = one smc_run() over all runs (run kernel + on-GPU reduction), the
model tables already resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

VALU_PEAK = 256 * 2.4e9  # wave64 VALU instructions per second: 256 CUs x 2.4 GHz x 1/clk/CU


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--runs", type=int, default=1_000_000)
    ap.add_argument("--processes", type=int, default=32)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(runs shard by run range; see DESIGN.md)", file=sys.stderr)
        return 2

    import cusmc_amd as cs
    from cusmc_amd import generators as G

    xml, _ = G.fischer_network(args.processes)
    net = cs.ParsedNetwork.from_string(xml, "P0.cs")
    q = cs.reach_query(300.0)
    sim = cs.GpuSimulator(net, device=0)
    for w in range(args.warmup):
        sim.run(q, seed=args.seed + 1000 + w, n_runs=args.runs)
    results = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        results.append(sim.run(q, seed=args.seed + k, n_runs=args.runs))
    wall = time.perf_counter() - t0
    sim.close()

    runs = sum(r.n_runs for r in results)
    trans = sum(r.transitions for r in results)
    steps = sum(r.steps for r in results)
    kms = sum(r.kernel_ms for r in results)
    value = runs / wall
    line = {
        "metric": "runs/sec (Fischer-32, Pr[<=300](<> P0.cs)); transitions/sec reported alongside",
        "value": value,
        "unit": "runs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Fischer-32 network generated as UPPAAL XML)",
        "config": {"workload": f"fischer-{args.processes} Pr[<=300](<> P0.cs)", "runs_per_step": args.runs,
                   "lanes_per_run": results[-1].lanes_per_run, "waves": results[-1].waves},
        "transitions_per_sec": trans / wall,
        "race_steps_per_sec": steps / wall,
        "kernel_runs_per_sec": runs / (kms / 1e3),
        "estimate": results[-1].estimate,
        "ch_eps": results[-1].ch_eps,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(net, q, args)
    print(json.dumps(line))
    return 0


def cpu_baseline(net, q, args) -> dict:
    """The oracle (cuSMC's per-component race, LITERAL mode) on the host
    cores, on a bounded sample of the same workload."""
    import oracle
    threads = os.cpu_count() or 1
    n = 256
    res, _ = oracle.run(net, q, seed=args.seed, n_runs=n, mode=oracle.LITERAL, threads=threads)
    rate = n / max(res.seconds, 1e-9)
    n = max(threads, int(rate * args.cpu_seconds))
    res, _ = oracle.run(net, q, seed=args.seed, n_runs=n, mode=oracle.LITERAL, threads=threads)
    return {"value": n / res.seconds, "unit": "runs/s", "cores": res.threads, "kind": "port",
            "sample": f"{n} runs of the same workload, oracle literal mode, {res.seconds:.1f} s"}


if __name__ == "__main__":
    sys.exit(main())
