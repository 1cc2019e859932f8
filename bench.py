"""Benchmark: statistical model checking on one MI355X.  Prints one JSON line.

Default workload: Fischer mutual exclusion with 32 processes, query
Pr[<=300](<> P0.cs), 1,000,000 runs per step, FP64 clocks and time.  One
step = one smc_run() over all runs (run kernel + on-GPU reduction); the model
tables are resident in HBM before the timed region, and the launch shape is
picked by the autotuner before it.  The default line also carries, under
`learned_is`, the rare-event headline of the `is` workload below (wall time
to +-10% with the learned sampler against plain simulation), measured in the
same process after the timed steps.

Other workloads (--workload):
  sir350   agent SIR, N=350, beta=0.3, gamma=0.1, Pr[<=100](<> I>=105), 1e6 runs
  race8    8 single-clock race components, Pr[<=100](<> count>=40), 1e4 runs
  batch    256 SIR + 256 Fischer instances of uneven size, 1e5 runs each,
           one graph-captured round per step
  is       rare-event SIR (N=100, beta=0.15, gamma=0.1, one infected),
           Pr[<=50](<> I>=90): wall time to a +-10% 95% interval with the
           learned importance sampler (training included) against the plain
           simulator
  copy     a 1 GiB device copy: the calibration of HBM byte counters

roofline: the run kernel is bound by VALU issue (tables sit in L2, per-run
output is a few bytes), so `achieved` counts algorithmic lane-operations:
the per-race-step figure of the oracle's cost model for the workload
(measured over 2e4+ runs, stated below) times the race steps of the launch,
divided by the run kernel's time from HIP events on its stream; the peak is
256 CUs x 64 lanes x 2.4 GHz.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

LANE_PEAK = 256 * 64 * 2.4e9        # VALU lane-operations per second
WAVE_PEAK = 256 * 2.4e9             # wave64 VALU instructions per second
# algorithmic lane-ops per race step (oracle cost model; tests pin them)
COST_PER_STEP = {"fischer32": 598.2, "sir350": 2214.6, "race8": 639.4}
SEED = 20261202
# Counter figures of the run kernel from rocprofv3 runs of this file (the
# commands and raw values are in profiles/), for the launch shape (lanes per
# run, register policy, table placement) they were taken on; another shape
# reports null.  traffic: HBM bytes per run (FETCH_SIZE + WRITE_SIZE, scaled by
# what they read for a 1 GiB device copy).  valu: SQ_INSTS_VALU (wave64 VALU
# instructions) per race step.
PROFILED = {
    "fischer32": {"shape": (4, 3, 1), "traffic_per_run": 23.1, "valu_per_step": 113.9,
                  "source": "profiles/sim_kernel_fischer32.md"},
    "sir350": {"shape": (64, 3, 2), "traffic_per_run": 25.9, "valu_per_step": 1543.6,
               "source": "profiles/sim_kernel_sir350.md"},
}


def _models(workload):
    import cusmc_amd as cs
    from cusmc_amd import generators as G
    if workload == "fischer32":
        net = cs.ParsedNetwork.from_string(G.fischer_network(32)[0], "P0.cs")
        return net, cs.reach_query(300.0), 1_000_000, "fischer-32 Pr[<=300](<> P0.cs)"
    if workload == "sir350":
        net = cs.ParsedNetwork.from_string(G.sir_network(350, 0.3, 0.1, seed=SEED)[0], "")
        return net, cs.var_query(net.variable_id("I"), 100.0, threshold=104), 1_000_000, \
            "sir-350 Pr[<=100](<> I>=105)"
    if workload == "race8":
        net = cs.ParsedNetwork.from_string(G.race_network(8, seed=SEED)[0], "")
        return net, cs.var_query(net.variable_id("count"), 100.0, threshold=39), 10_000, \
            "race-8 Pr[<=100](<> count>=40)"
    raise ValueError(workload)


def bench_plain(args) -> dict:
    import cusmc_amd as cs
    net, q, runs, name = _models(args.workload)
    runs = args.runs or runs
    sim = cs.GpuSimulator(net, device=0)
    if args.lanes:
        # a fixed launch shape (profiles pin one kernel build this way)
        shape = {"lanes_per_run": args.lanes, "occupancy": args.occupancy, "tables": args.tables}
        policy, tables = args.occupancy, args.tables
    else:
        tuned = sim.autotune(q, seed=SEED + 999, sample_runs=max(1024, min(runs // 4, 16384)))
        shape, policy, tables = {}, tuned.occupancy, tuned.tables
    for w in range(args.warmup):
        sim.run(q, seed=SEED + 1000 + w, n_runs=runs, **shape)
        _progress(f"warmup {w + 1}/{args.warmup}")
    results = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        results.append(sim.run(q, seed=SEED + k, n_runs=runs, **shape))
        _progress(f"step {k + 1}/{args.steps}: {results[-1].kernel_ms:.1f} ms kernel")
    wall = time.perf_counter() - t0
    sim.close()
    n = sum(r.n_runs for r in results)
    trans = sum(r.transitions for r in results)
    steps = sum(r.steps for r in results)
    kms = sum(r.kernel_ms for r in results)
    achieved = COST_PER_STEP[args.workload] * steps / (kms / 1e3)
    used = (results[-1].lanes_per_run, policy, int(tables))
    prof = PROFILED.get(args.workload, {})
    same_shape = prof.get("shape") == used
    traffic = prof["traffic_per_run"] * runs if same_shape and prof.get("traffic_per_run") else None
    valu_frac = (prof["valu_per_step"] * steps / (kms / 1e3) / WAVE_PEAK
                 if same_shape and prof.get("valu_per_step") else None)
    line = {
        "metric": "runs/sec; transitions/sec reported alongside",
        "value": n / wall,
        "unit": "runs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded network generated as UPPAAL XML)",
        "config": {"workload": name, "runs_per_step": runs, "lanes_per_run": results[-1].lanes_per_run,
                   "waves": results[-1].waves, "occupancy_policy": policy, "tables_in_lds": bool(tables),
                   "table_mode": int(tables),
                   "launch": "fixed" if args.lanes else "autotuned"},
        "transitions_per_sec": trans / wall,
        "race_steps_per_sec": steps / wall,
        "kernel_runs_per_sec": n / (kms / 1e3),
        "kernel_ms_per_step": kms / args.steps,
        "estimate": results[-1].estimate,
        "ch_eps": results[-1].ch_eps,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": LANE_PEAK, "unit": "lane-op/s",
                     "frac": achieved / LANE_PEAK, "traffic": traffic,
                     "work": f"{COST_PER_STEP[args.workload]} algorithmic lane-ops per race step x "
                             f"{steps // args.steps} race steps per launch",
                     # issued wave64 VALU instructions / (kernel time x 256 CUs x 2.4 GHz)
                     "valu_issue_frac": valu_frac,
                     "counters_source": (prof.get("source") + " (rocprofv3 on this launch shape)")
                     if same_shape and (traffic or valu_frac) else None},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(net, q, args)
    if args.workload == "fischer32" and not args.no_is:
        # the learned importance sampler's headline (rare-event SIR-100)
        _progress("learned IS leg")
        r = bench_is(args)
        line["learned_is"] = {k: r[k] for k in (
            "is_seconds", "is_train_seconds", "is_estimate", "is_rel_half_width", "plain_seconds_to_ci",
            "plain_estimate", "plain_rel_half_width", "plain_runs_per_sec", "plain_extrapolated",
            "is_inside_plain_interval")}
        line["learned_is"]["speedup"] = r["value"]
        line["learned_is"]["workload"] = r["config"]["workload"]
        line["learned_is"]["exact"] = r["config"]["exact"]
    return line


def bench_batch(args) -> dict:
    import cusmc_amd as cs
    from cusmc_amd import generators as G
    runs = args.runs or 100_000
    specs = G.batch_instances(seed=SEED)
    sims, items = [], []
    for k, sp in enumerate(specs):
        net = cs.ParsedNetwork.from_string(sp["xml"], sp["loc_query"])
        sim = cs.GpuSimulator(net, device=0)
        if sp["kind"] == "sir":
            q = cs.var_query(net.variable_id(sp["var"]), sp["time_bound"], threshold=sp["threshold"])
        else:
            q = cs.reach_query(sp["time_bound"])
        sims.append(sim)
        items.append((sim, q, SEED + k, runs))
    t_build = time.perf_counter()
    b = cs.Batch(items, tune=True)
    t_build = time.perf_counter() - t_build
    _progress(f"batch of {len(items)} instances captured in {t_build:.1f} s")
    for _ in range(args.warmup):
        b.run(1)
        _progress("warmup round")
    t0 = time.perf_counter()
    gms = 0.0
    tot_runs = tot_trans = tot_steps = 0
    for _ in range(args.steps):
        res, st = b.run(1)
        gms += st.graph_ms
        tot_runs += st.runs
        tot_trans += st.transitions
        tot_steps += st.steps
        _progress(f"round: {st.graph_ms:.1f} ms")
    wall = time.perf_counter() - t0
    groups = st.groups
    b.close()
    for s in sims:
        s.close()
    return {
        "metric": "runs/sec over a batch of uneven instances; transitions/sec reported alongside",
        "value": tot_runs / wall, "unit": "runs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (256 SIR + 256 Fischer networks generated from the seed)",
        "config": {"workload": "batch 256 sir + 256 fischer", "runs_per_instance": runs, "groups": groups,
                   "tune_and_capture_s": t_build},
        "transitions_per_sec": tot_trans / wall, "race_steps_per_sec": tot_steps / wall,
        "graph_ms_per_step": gms / args.steps,
    }


def bench_is(args) -> dict:
    """Time to +-10% at 95%: learned IS (training included) vs plain."""
    import cusmc_amd as cs
    from cusmc_amd import generators as G
    net = cs.ParsedNetwork.from_string(G.sir_network(100, 0.15, 0.1, initial_infected=1, seed=SEED)[0], "")
    q = cs.var_query(net.variable_id("I"), 50.0, threshold=89)
    exact = 6.638e-6  # exact lumped CTMC of this network and query
    sim = cs.GpuSimulator(net, device=0)
    sim.autotune(q, seed=SEED + 999, sample_runs=65536)
    sim.run(q, seed=SEED + 998, n_runs=1 << 20)  # warm
    # plain: batches of runs until the CLT half-width is within 10%, capped
    # (a cap of 0 skips it: the learned sampler alone)
    t0 = time.perf_counter()
    hits = n = 0
    batch = args.runs or 8_000_000
    rel = math.inf
    while args.plain_cap_s > 0 and time.perf_counter() - t0 < args.plain_cap_s:
        r = sim.run(q, seed=SEED, n_runs=batch, run_begin=n)
        hits += r.hits
        n += r.n_runs
        p = hits / n
        rel = 1.959963984540054 * math.sqrt(p * (1 - p) / n) / p if hits else math.inf
        print(f"plain: {n} runs, {hits} hits, rel half-width {rel:.3f}", file=sys.stderr, flush=True)
        if hits >= 10 and rel <= 0.1:
            break
    plain_s = time.perf_counter() - t0
    plain_p = hits / n if n else math.nan
    plain_rate = n / plain_s if n else math.nan
    # runs plain needs for +-10% at the exact p, and the time at its rate
    need = (1.959963984540054 / 0.1) ** 2 * (1 - exact) / exact
    plain_to_ci_s = plain_s if rel <= 0.1 else (need / plain_rate if n else math.nan)
    # learned IS: create, train (--is-rounds x --is-runs), estimate to +-10%
    t1 = time.perf_counter()
    s = cs.ImportanceSampler(sim, q, rounds=args.is_rounds, runs_per_round=args.is_runs, epochs=args.epochs,
                             learning_rate=args.lr, seed=SEED)
    rounds = s.train()
    t_train = time.perf_counter() - t1
    est = s.estimate(seed=SEED + 1, batch_runs=65536, max_runs=1 << 24, rel_target=0.1)
    is_s = time.perf_counter() - t1
    s.close()
    sim.close()
    half = est.half_width
    plain_half = 1.959963984540054 * math.sqrt(plain_p * (1 - plain_p) / n) if hits else math.inf
    return {
        "metric": "time to +-10% relative half-width at 95%, learned IS vs plain (speed-up)",
        "value": plain_to_ci_s / is_s if n else None, "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": is_s * 1e3, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64 (simulation), bf16/fp32 (MLP)",
        "data": "synthetic (SIR N=100 network generated from the seed)",
        "config": {"workload": "is sir-100 Pr[<=50](<> I>=90)", "exact": exact, "rounds": args.is_rounds,
                   "runs_per_round": args.is_runs, "epochs": args.epochs, "learning_rate": args.lr},
        "is_seconds": is_s, "is_train_seconds": t_train, "is_estimate": est.estimate,
        "is_rel_half_width": est.rel_half_width, "is_runs": est.n_runs, "is_hits": est.hits, "is_ess": est.ess,
        "plain_seconds_to_ci": plain_to_ci_s, "plain_measured_seconds": plain_s, "plain_runs": n,
        "plain_hits": hits, "plain_estimate": plain_p, "plain_rel_half_width": rel,
        "plain_runs_per_sec": plain_rate, "plain_extrapolated": rel > 0.1,
        "is_inside_plain_interval": bool(hits) and abs(est.estimate - plain_p) <= plain_half + half,
        "rounds": rounds,
    }


def _progress(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(net, q, args) -> dict:
    """The oracle on the host cores, on a bounded sample of the same
    workload: cuSMC's per-component race (literal mode: the reported
    baseline) on all cores the process may use and on one core, and beside
    it the GPU's own sampling of the race (canonical mode) on all cores, so
    the GPU/CPU ratio can be read with the algorithm held fixed."""
    import oracle

    def timed(threads, seconds, mode):
        n = max(threads, 64)
        res, _ = oracle.run(net, q, seed=SEED, n_runs=n, mode=mode, threads=threads)
        rate = n / max(res.seconds, 1e-9)
        n = max(threads, int(rate * seconds))
        res, _ = oracle.run(net, q, seed=SEED, n_runs=n, mode=mode, threads=threads)
        return n, res
    threads = _host_cores()
    n, res = timed(threads, args.cpu_seconds, oracle.LITERAL)
    n1, res1 = timed(1, max(2.0, args.cpu_seconds / 3), oracle.LITERAL)
    nc, resc = timed(threads, max(2.0, args.cpu_seconds / 3), oracle.CANONICAL)
    _progress(f"cpu baseline: {n / res.seconds:.1f} runs/s on {res.threads} threads "
              f"(canonical sampling {nc / resc.seconds:.1f} runs/s)")
    return {"value": n / res.seconds, "unit": "runs/s", "cores": res.threads, "kind": "port",
            "sample": f"{n} runs of the same workload, oracle literal mode, {res.seconds:.1f} s on "
                      f"{res.threads} threads ({_cpu_model()})",
            "single_core": {"value": n1 / res1.seconds, "unit": "runs/s", "sample": f"{n1} runs, {res1.seconds:.1f} s"},
            "canonical": {"value": nc / resc.seconds, "unit": "runs/s", "cores": resc.threads,
                          "sample": f"{nc} runs, oracle canonical mode (the GPU's sampling of the race), "
                                    f"{resc.seconds:.1f} s"}}


def bench_copy(args) -> dict:
    """A plain device copy of a known size: the calibration kernel for the
    HBM byte counters (FETCH_SIZE, WRITE_SIZE) of the profiles."""
    import torch
    nbytes = 1 << 30
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"metric": "device copy bandwidth (read + write)", "value": 2 * nbytes * args.steps / wall / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"copy {nbytes} bytes"}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="fischer32",
                    choices=["fischer32", "sir350", "race8", "batch", "is", "copy"])
    ap.add_argument("--runs", type=int, default=0, help="runs per step (default: the workload's)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--plain-cap-s", type=float, default=120.0, help="is: cap on the plain run's wall time")
    ap.add_argument("--epochs", type=int, default=6, help="is: epochs per cross-entropy round")
    ap.add_argument("--lr", type=float, default=1e-2, help="is: Adam step")
    ap.add_argument("--is-rounds", type=int, default=4, help="is: cross-entropy rounds")
    ap.add_argument("--is-runs", type=int, default=50_000, help="is: runs per cross-entropy round")
    ap.add_argument("--lanes", type=int, default=0, help="fixed lanes per run (default: autotuned)")
    ap.add_argument("--occupancy", type=int, default=0, choices=[0, 3],
                    help="with --lanes: register policy (0 unrestricted, 3 three waves per SIMD)")
    ap.add_argument("--tables", type=int, default=0, choices=[0, 1, 2],
                    help="with --lanes: model tables in global memory (0), staged in LDS (1), or staged "
                         "but for the location records (2)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-is", action="store_true", help="fischer32: skip the learned-IS leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not run here (the run range would be split "
              f"across GPUs, one process each, with a reduction of a few scalars)", file=sys.stderr)
        return 2
    if args.workload == "copy":
        line = bench_copy(args)
    elif args.workload == "batch":
        line = bench_batch(args)
    elif args.workload == "is":
        line = bench_is(args)
    else:
        line = bench_plain(args)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
