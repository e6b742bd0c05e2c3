#!/usr/bin/env python3
"""Benchmark of the clause-vs-assignment sweep on one MI355X.

Workload (the headline configuration): 1M variables, 4M clauses of length
2 + Geometric(0.1) capped at 200 (~44M literals, uint32 2*var+sign), 1024
partial assignments each assigning a variable with probability 0.4 (polarity
Bernoulli(0.5)), 0.1% planted hits; synthetic, seeded.

A step is one sweep: every clause against all 1024 assignments, hits compacted
on the device (gss_sweep_async + gss_sweep_wait).  Inputs are resident in HBM
before timing starts.  Prints one JSON line.

  value          literal x assignment checks per second (checks = L * A, no
                 early-exit credit) over the timed wall-clock region
  roofline       algorithmic bytes of the sweep kernel per launch
                 (4 L + 16 W V_used + 24 R) / its average duration from HIP
                 events recorded on the engine's stream, against 8 TB/s
  cpu_baseline   the CPU oracle (bit-sliced rule, all host threads) on a
                 bounded sample of the same instance
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("lit×assig checks/s + %HBM roofline @4M cl×1024 asg; p99 report latency")
# HBM bytes per sweep launch of the default kernel on this workload, from two
# rocprofv3 PMC passes (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, KiB, the
# factors calibrated on the clause-upload copy kernel of known size); the
# commands and raw figures are in profiles/k_sweep_sweep1024.md.  None when
# the instance is not the headline one.
MEASURED_TRAFFIC_KIB = {("k_sweep<16,1,4>", 1.0): 2 * 2_074_122 + 427_740}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(inst, target_s: float = 12.0) -> dict:
    """Times the oracle (a port of the reference rule) on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    nthreads = max(1, min(16, os.cpu_count() or 1))
    off = inst.offsets
    # bounded sample: a prefix of the clauses, all 1024 assignments
    n_sample = min(inst.lens.size, 1_000_000)
    sub_off = off[: n_sample + 1]
    lits = inst.lits[: sub_off[-1]]
    checks_per_run = int(sub_off[-1]) * inst.assignments
    t0 = time.perf_counter()
    runs = 0
    while True:
        orc.sweep(sub_off, lits, inst.is_def, inst.is_true, inst.valid, nthreads=nthreads)
        runs += 1
        el = time.perf_counter() - t0
        if el >= target_s or runs >= 50:
            break
    all_cores = checks_per_run * runs / el
    # one thread, smaller prefix
    n1 = min(n_sample, 100_000)
    t0 = time.perf_counter()
    orc.sweep(off[: n1 + 1], inst.lits[: off[n1]], inst.is_def, inst.is_true, inst.valid,
              nthreads=1)
    one = int(off[n1]) * inst.assignments / (time.perf_counter() - t0)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": all_cores, "unit": "checks/s", "cores": nthreads, "kind": "port",
            "sample": f"first {n_sample} clauses ({int(sub_off[-1])} literals) x "
                      f"{inst.assignments} assignments, {runs} runs, {nthreads} threads",
            "value_1core": one, "cpu_model": model}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0, help="instance scale (1.0 = headline)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--words-tile", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the sweep is built and measured on a single "
              "MI355X (see DESIGN.md for how it would shard); run with --gpus 1",
              file=sys.stderr)
        return 2

    from gpusharesat_amd import workloads
    from gpusharesat_amd.engine import Engine

    t0 = time.time()
    log("generating workload sweep-1024 ...")
    inst = workloads.sweep_1024(seed=2, scale=args.scale)
    L, A = int(inst.lits.size), inst.assignments
    v_used = inst.used_vars
    log(f"generated in {time.time() - t0:.1f}s: {inst.var_count} vars, {inst.lens.size} clauses, "
        f"{L} literals, {A} assignments")

    e = Engine(0)
    e.append(inst.lits, inst.lens)
    e.reset_assignments(inst.var_count, inst.words)
    e.upload(inst.is_def, inst.is_true)
    valid = inst.valid
    log("uploaded; warming up")
    hits = 0
    for _ in range(args.warmup):
        e.sweep_async(valid)
        hits = e.sweep_wait()
    e.sync()
    e.kernel_stats(reset=True)
    log(f"timing {args.steps} steps")
    t_start = time.perf_counter()
    for _ in range(args.steps):
        e.sweep_async(valid)
        hits = e.sweep_wait()
    e.sync()
    elapsed = time.perf_counter() - t_start
    k_total_ms, k_count = e.kernel_stats()
    k_avg_s = k_total_ms / 1000.0 / max(1, k_count)
    ms_per_step = elapsed * 1000.0 / args.steps
    checks = L * A
    value = checks * args.steps / elapsed
    W = inst.words
    alg_bytes = 4 * L + 16 * W * v_used + 24 * hits
    achieved = alg_bytes / k_avg_s / 1e9
    log(f"{ms_per_step:.3f} ms/step, kernel {k_avg_s * 1e3:.3f} ms, {hits} hit records, "
        f"{achieved:.0f} GB/s algorithmic")
    cpu = None
    if not args.no_cpu:
        log("timing the CPU baseline (oracle) ...")
        cpu = cpu_baseline(inst)
        log(f"cpu: {cpu['value']:.3e} checks/s on {cpu['cores']} threads")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "sweep-1024: 1M vars, 4M clauses len 2+Geom(0.1) cap 200, "
                        "1024 assignments 40% assigned, 0.1% planted",
            "vars": inst.var_count, "clauses": int(inst.lens.size), "literals": L,
            "assignments": A, "words": W, "used_vars": v_used, "hit_records": int(hits),
            "scale": args.scale,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (MEASURED_TRAFFIC_KIB[("k_sweep<16,1,4>", args.scale)] * 1024
                        if ("k_sweep<16,1,4>", args.scale) in MEASURED_TRAFFIC_KIB else None),
            "traffic_unit": "bytes per launch",
            "kernel": "k_sweep<16,1,4>",
            "kernel_ms": k_avg_s * 1e3,
            "algorithmic_bytes": alg_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    e.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
