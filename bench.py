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
                 events recorded on the engine's stream, against 8 TB/s;
                 beside it the kernel's traffic past the L2 (two PMC passes)
                 as a rate, and the device's random 128-byte line gather rate
                 on a table of the same size, measured in the same run
  cpu_baseline   the CPU oracle (bit-sliced rule) on the whole instance, with
                 the threads of this process's CPU share (the box's share
                 for one GPU); nproc and the CPU model are recorded beside it
  latency_ms     after the timed region: submit -> every hit of the sweep in
                 pinned host memory as the drop-in gpuRun() delivers it (one
                 16-byte record per clause with its solver words), one sweep
                 in flight, through the step path (graph replay, device
                 compaction, PCIe copy); latency_id_records_ms the same with
                 one 8-byte record per (clause, word), latency_full_records_ms
                 with the 24-byte records that also carry the masks

--workload batched-16 measures the batched form instead: 16 solver streams
over 1M vars, each step inserting Poisson(16 x 2800) clauses of Pareto(1.5)
lengths (2..500) and 16 uneven batches of U[16, 256] assignments (64-word
table, 4096 slots), as one graph-captured step (gss_step_enqueue /
gss_step_wait / gss_step_hits_copy, two steps in flight).  The DB grows from
1M to ~10M clauses over 200 steps; the warmup steps are also replayed through
the plain entry points on a second engine and their records must be
identical.  value = sum over timed steps of (DB literals x valid assignments)
/ wall-clock; latency = enqueue -> records on the host, per step.
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
# HBM bytes per sweep launch of the default kernel on the headline instance
# come from two rocprofv3 PMC passes (FETCH_SIZE, WRITE_SIZE) recorded in this
# profile, together with the SHA-256 of the kernel source they were measured
# on; a profile of another source is not used (traffic is then null).
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_sweep_sweep1024.md")
KERNEL_SOURCE = os.path.join(ROOT, "gpusharesat_amd", "csrc", "kernels.hip")


def measured_traffic(kernel: str) -> dict:
    """Traffic of `kernel` on the headline instance from the committed
    profile, if that profile was taken on the current kernel source."""
    import hashlib
    import re
    src = hashlib.sha256(open(KERNEL_SOURCE, "rb").read()).hexdigest()
    out = {"bytes": None, "profile": os.path.relpath(TRAFFIC_PROFILE, ROOT),
           "kernel_source_sha256": src}
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        out["note"] = "no traffic profile"
        return out
    m = re.search(r"traffic-record: kernel=(\S+) source_sha256=([0-9a-f]{64}) "
                  r"fetch_kib=(\d+) fetch_factor=([\d.]+) write_kib=(\d+) "
                  r"write_factor=([\d.]+) date=(\S+)", text)
    if not m:
        out["note"] = "the profile holds no traffic record"
        return out
    k, sha, fk, ff, wk, wf, date = m.groups()
    out.update({"profiled_kernel": k, "profiled_source_sha256": sha, "date": date})
    if k != kernel or sha != src:
        out["note"] = "the profile was taken on another kernel or kernel source"
        return out
    out["bytes"] = int((int(fk) * float(ff) + int(wk) * float(wf)) * 1024)
    return out


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_share() -> tuple[int, dict]:
    """Threads of this process's CPU share, and what the host reports.  The
    share is the CPUs this process may run on, bounded by OMP_NUM_THREADS
    where the environment sets it (the box gives one GPU's process 16)."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    share = affinity
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        share = min(share, int(env))
    info = {"nproc": nproc, "affinity_cpus": affinity, "omp_num_threads": env or None}
    try:
        fields = {}
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "vendor_id", "cpu family", "model", "stepping",
                         "cpu MHz", "cache size") and k not in fields:
                    fields[k] = v
        info["cpu_model"] = fields.get("model name", "")
        info["cpuinfo"] = fields
    except OSError:
        pass
    return max(1, share), info


def cpu_baseline(inst, target_s: float = 10.0) -> dict:
    """Times the oracle (a port of the reference rule, bit-sliced, 64
    assignments per word) on the whole instance with every thread of the
    process's CPU share, repeated for about target_s; one thread is timed on
    the first 100k clauses."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as orc
    nthreads, info = cpu_share()
    off = inst.offsets
    checks_per_run = int(off[-1]) * inst.assignments
    t0 = time.perf_counter()
    runs = 0
    while True:
        orc.sweep(off, inst.lits, inst.is_def, inst.is_true, inst.valid, nthreads=nthreads)
        runs += 1
        el = time.perf_counter() - t0
        if el >= target_s or runs >= 50:
            break
    all_cores = checks_per_run * runs / el
    n1 = min(inst.lens.size, 100_000)
    t0 = time.perf_counter()
    orc.sweep(off[: n1 + 1], inst.lits[: off[n1]], inst.is_def, inst.is_true, inst.valid,
              nthreads=1)
    one = int(off[n1]) * inst.assignments / (time.perf_counter() - t0)
    return {"value": all_cores, "unit": "checks/s", "cores": nthreads, "kind": "port",
            "sample": f"the whole instance ({inst.lens.size} clauses, {int(off[-1])} literals) "
                      f"x {inst.assignments} assignments, {runs} runs in {el:.1f} s, "
                      f"{nthreads} threads",
            "value_1core": one, "sample_1core": f"first {n1} clauses, 1 thread",
            "seconds_per_sweep": el / runs, **info}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (sweep-1024: 50, batched-16: 196, perftest: 2000)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps (sweep-1024: 5, batched-16: 4, perftest: 20)")
    ap.add_argument("--scale", type=float, default=1.0, help="instance scale (1.0 = headline)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--inflight", type=int, default=2, choices=(1, 2),
                    help="batched-16, serving-32: steps on the device at once "
                         "(2 = throughput, 1 = latency)")
    ap.add_argument("--plain-steps", type=int, default=20,
                    help="batched-16: steps also timed through the plain entry points")
    ap.add_argument("--lat-steps", type=int, default=30,
                    help="sweep-1024: sweeps timed for report latency after the timed steps")
    ap.add_argument("--workload", default="sweep-1024",
                    choices=["sweep-1024", "batched-16", "serving-32", "residency", "ceilings",
                             "perftest"])
    ap.add_argument("--duration", type=float, default=60.0, help="serving-32: seconds")
    ap.add_argument("--rate", type=float, default=2000.0, help="serving-32: requests per second")
    ap.add_argument("--preload", type=int, default=0,
                    help="serving-32: clauses in the DB before the run (default none)")
    ap.add_argument("--screen", action="store_true",
                    help="batched-16: steps submitted with GSS_STEP_SCREEN (clauses screened "
                         "against the table's aggregate before the exact check)")
    ap.add_argument("--flips", type=int, default=32,
                    help="serving-32: variables each assignment redraws from its solver's base")
    ap.add_argument("--records", default="full", choices=["full", "ids", "clauses"],
                    help="batched-16, serving-32: records with masks (full), clause and word "
                         "only (ids), or (batched-16) one record per clause with its words "
                         "(clauses, what gpuRun delivers)")
    ap.add_argument("--signs", default="planted", choices=["planted", "random"],
                    help="serving-32, batched-16: clauses and assignments near a planted "
                         "assignment, or random signs and polarities")
    ap.add_argument("--res-scale", type=float, default=1.0, help="residency: workload scale")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched-16, serving-32: default tile settings")
    ap.add_argument("--tuning", default=None,
                    help="batched-16, serving-32: fixed settings instead of the autotuner, "
                         "e.g. unroll=2,fuse_len=-1,task_clauses=32,waves_per_cu=0,schedule=1")
    args = ap.parse_args()
    if args.tuning:
        args.tuning = {k: int(v) for k, v in
                       (item.split("=", 1) for item in args.tuning.split(","))}
    batched = args.workload == "batched-16"
    if args.steps is None:
        args.steps = {"batched-16": 196, "perftest": 2000}.get(args.workload, 50)
    if args.warmup is None:
        args.warmup = {"batched-16": 4, "perftest": 20}.get(args.workload, 5)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the sweep is built and measured on a single "
              "MI355X (a multi-GPU run would shard the clause DB, not the slots); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "batched-16":
        return run_batched(args)
    if args.workload == "serving-32":
        return run_serving(args)
    if args.workload == "residency":
        return run_residency(args)
    if args.workload == "ceilings":
        return run_ceilings(args)
    if args.workload == "perftest":
        return run_perftest(args)

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
    # Report latency (outside the timed region): submit -> every hit of the
    # sweep in pinned host memory, one sweep in flight, through the step path
    # (graph replay: sweep, compaction, copy over PCIe); first with the
    # records the solvers receive (clause and solver word), then with the
    # full records.
    from gpusharesat_amd import serving as sv
    views = {}
    sizes = {sv.RECORDS_CLAUSES: 16, sv.RECORDS_IDS: 8, sv.RECORDS_FULL: 24}

    def latency_of(records: int) -> dict | None:
        lat = []
        for k in range(args.lat_steps + 2):
            t_sub = time.perf_counter()
            sv.submit(e, valid, records=records)
            n_step, _ = e.step_wait()
            view = (e.step_clauses_view() if records == sv.RECORDS_CLAUSES else
                    e.step_ids_view() if records == sv.RECORDS_IDS else e.step_hits_view())
            t_rep = time.perf_counter()
            if view.size != n_step or (records != sv.RECORDS_CLAUSES and n_step != hits):
                raise RuntimeError(f"step path gave {n_step} records, the sweep {hits}")
            if records not in views:
                views[records] = view.copy()
            if k >= 2:  # the first replays capture the graph and size the buffers
                lat.append((t_rep - t_sub) * 1e3)
        if not lat:
            return None
        return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                "p999": float(np.percentile(lat, 99.9)), "max": float(max(lat)),
                "sweeps": len(lat), "records": int(views[records].size),
                "bytes_per_record": sizes[records],
                "what": "submit -> every hit record of the sweep in pinned host memory"}

    latency = latency_ids = latency_full = None
    if args.lat_steps > 0:
        latency = latency_of(sv.RECORDS_CLAUSES)
        latency_ids = latency_of(sv.RECORDS_IDS)
        latency_full = latency_of(sv.RECORDS_FULL)
        log(f"report latency, one 16-byte record per clause ({latency['records']}): p50 "
            f"{latency['p50']:.2f} ms, p99 {latency['p99']:.2f} ms; 8-byte per (clause, word) "
            f"p99 {latency_ids['p99']:.2f} ms; 24-byte p99 {latency_full['p99']:.2f} ms")
        # the three forms carry the same reports: ids = the full records'
        # first words; clause records = the ids grouped by clause
        ids = views[sv.RECORDS_IDS]
        full = views[sv.RECORDS_FULL]
        if not np.array_equal(np.sort(ids.view(np.uint64)),
                              np.sort(full.view(np.uint64).reshape(-1, 3)[:, 0])):
            raise RuntimeError("id records differ from the full records' first words")
        key = ids["row"].astype(np.uint64) | (ids["len"].astype(np.uint64) << np.uint64(32))
        uk, inv = np.unique(key, return_inverse=True)
        words = np.zeros(uk.size, np.uint64)
        np.bitwise_or.at(words, inv, np.uint64(1) << ids["word"].astype(np.uint64))
        cl = views[sv.RECORDS_CLAUSES]
        ck = cl["row"].astype(np.uint64) | (cl["len"].astype(np.uint64) << np.uint64(32))
        o = np.argsort(ck)
        if not (np.array_equal(ck[o], uk) and np.array_equal(cl["words"][o], words)):
            raise RuntimeError("clause records differ from the id records grouped by clause")
    # The device's own ceiling for this access pattern: random 128-byte line
    # gathers (independent, 4 in flight per lane) from a table of the same size.
    kernel_name = "k_sweep<16,1,4,true>"
    tr = measured_traffic(kernel_name) if args.scale == 1.0 else {
        "bytes": None, "note": "not the headline instance"}
    traffic = tr["bytes"]
    table_bytes = (2 * inst.var_count + 2) * W * 8
    gather = _gather_ceiling((table_bytes + 127) // 128 * 128)
    log(f"random-line gather ceiling on a {table_bytes / 1e6:.0f} MB table: {gather:.0f} GB/s"
        + (f"; the sweep's measured traffic runs at {traffic / k_avg_s / 1e9:.0f} GB/s"
           if traffic else ""))
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
            "traffic": traffic,
            "traffic_unit": "bytes per launch",
            "traffic_source": tr,
            "traffic_gbps": traffic / k_avg_s / 1e9 if traffic else None,
            "gather_ceiling_gbps": gather,
            "kernel": "k_sweep<16,1,4,true>",
            "kernel_ms": k_avg_s * 1e3,
            "algorithmic_bytes": alg_bytes,
        },
        "cpu_baseline": cpu,
        "latency_ms": latency,
        "latency_id_records_ms": latency_ids,
        "latency_full_records_ms": latency_full,
        "step_hit_buffer": e.step_stats(),
    }
    print(json.dumps(out), flush=True)
    e.close()
    return 0


def _gather_ceiling(table_bytes: int) -> float:
    """GB/s of random 128-byte line reads (4 in flight per lane) from a
    device table of table_bytes (gss_ceiling_gather)."""
    import ctypes
    from gpusharesat_amd import _lib
    gbps, ms = ctypes.c_double(), ctypes.c_double()
    _lib.call("gss_ceiling_gather", 0, table_bytes, 32 << 20, 4, 5, ctypes.byref(gbps),
              ctypes.byref(ms))
    return gbps.value


def _sorted_hits(h: np.ndarray) -> np.ndarray:
    return np.sort(h, order=["len", "row", "word"])


def run_batched(args) -> int:
    from gpusharesat_amd import workloads
    from gpusharesat_amd.engine import Engine

    steps_total = args.warmup + args.steps
    t0 = time.time()
    log("generating workload batched-16 ...")
    # random signs: clause signs and assigned values uniform, 40 % of the
    # variables assigned (as the headline sweep); planted: near a solution
    gen = workloads.SolverStreams(seed=3, var_count=1_000_000,
                                  **({} if args.signs == "planted" else
                                     dict(p_assigned=0.4, agree=0.5, lit_true=0.5)))
    d0, t0p = gen.initial_table()
    init_lits, init_lens = gen.initial_clauses(1_000_000)
    plan = [gen.step() for _ in range(steps_total)]
    log(f"generated {steps_total} steps in {time.time() - t0:.1f}s; "
        f"{sum(p.lens.size for p in plan)} clauses inserted, "
        f"{sum(p.ups.size for p in plan)} updates")

    e = Engine(0)
    e.append(init_lits, init_lens)
    e.reset_assignments(gen.V, gen.words)
    e.upload(d0, t0p)
    tuned = None
    if args.tuning:
        e.apply_tuning(args.tuning)
        tuned = args.tuning
    elif not args.no_autotune:
        tuned, tms = e.autotune(plan[0].valid)
        log(f"autotuned: {tuned} ({tms:.3f} ms per sweep of the initial DB)")
    # the warmup steps are checked against the plain entry points
    ref = Engine(0)
    ref.append(init_lits, init_lens)
    ref.reset_assignments(gen.V, gen.words)
    ref.upload(d0, t0p)
    if tuned:
        ref.apply_tuning(tuned)
    del d0, t0p
    verified = 0
    for k in range(args.warmup):
        st = plan[k]
        e.step_enqueue(st.lits, st.lens, st.ups, st.valid, st.last_pos, screen=args.screen)
        n, _ = e.step_wait()
        got = _sorted_hits(e.step_hits())
        ref.append(st.lits, st.lens)
        ref.apply(st.ups)
        ref.sweep(st.valid)
        exp = _sorted_hits(ref.hits())
        ref.roll_forward(st.ups, st.last_pos)
        if got.size != n or not np.array_equal(got, exp):
            log(f"step {k}: batched records differ from the plain replay")
            return 1
        verified += 1
    # the plain form on the steps the batched run times first: one step at a
    # time through the plain entry points, records copied to the host
    n_plain = min(args.plain_steps, args.steps)
    plain_checks = []
    lits_ref = ref.totals()[1]
    ref.sync()
    t_plain = time.perf_counter()
    for k in range(args.warmup, args.warmup + n_plain):
        st = plan[k]
        ref.append(st.lits, st.lens)
        ref.apply(st.ups)
        ref.sweep(st.valid)
        ref.hits()
        ref.roll_forward(st.ups, st.last_pos)
        lits_ref += int(st.lits.size)
        plain_checks.append(lits_ref * st.assignments)
    ref.sync()
    plain_s = time.perf_counter() - t_plain
    ref.close()
    log(f"{verified} warmup steps identical to the plain replay; plain form "
        f"{sum(plain_checks) / plain_s:.3e} checks/s over {n_plain} steps; "
        f"timing {args.steps} steps")
    lits_db = e.totals()[1]
    checks = 0
    lat = []
    step_ms = []
    done_t = []
    hits_total = 0
    pending = []
    e.sync()
    t_start = time.perf_counter()
    todo = list(range(args.warmup, steps_total))

    from gpusharesat_amd import serving as sv
    rec = {"ids": sv.RECORDS_IDS, "clauses": sv.RECORDS_CLAUSES}.get(args.records,
                                                                    sv.RECORDS_FULL)

    submit_s = []

    def enqueue_next():
        nonlocal lits_db, checks
        st = plan[todo.pop(0)]
        t = time.perf_counter()
        pending.append(t)
        sv.submit(e, st.valid, lits=st.lits, lens=st.lens, ups=st.ups, last_pos=st.last_pos,
                  records=rec, flags=sv.STEP_SCREEN if args.screen else 0)
        submit_s.append(time.perf_counter() - t)
        lits_db += int(st.lits.size)
        checks += lits_db * st.assignments

    # two steps in flight (default); once the older one ends its records
    # travel to the host while the next step is staged, then they are read.
    # With --inflight 1 a step is submitted only once the previous one has
    # left the device (its records still travelling), so no step waits
    # behind another's graph.
    while todo and len(pending) < args.inflight:
        enqueue_next()
    while pending:
        _, ms = e.step_wait()
        if todo:
            enqueue_next()
        # records on the host (pinned, no copy)
        hits_total += (e.step_ids_view() if rec == sv.RECORDS_IDS else
                       e.step_clauses_view() if rec == sv.RECORDS_CLAUSES else
                       e.step_hits_view()).size
        done_t.append(time.perf_counter())
        lat.append(done_t[-1] - pending.pop(0))
        step_ms.append(ms)
    e.sync()
    elapsed = time.perf_counter() - t_start
    clauses_end = e.totals()[0]
    lat_ms = np.array(lat) * 1e3
    value = checks / elapsed
    log(f"{elapsed * 1e3 / args.steps:.3f} ms/step, {value:.3e} checks/s, "
        f"p50 {np.percentile(lat_ms, 50):.2f} ms, p99 {np.percentile(lat_ms, 99):.2f} ms, "
        f"DB {clauses_end} clauses")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "batched-16: 1M vars, 16 streams x U[16,256] assignments (64 words), "
                        "Poisson(16x2800) Pareto(1.5) clauses per step, DB 1M -> ~10M",
            "clauses_end": int(clauses_end), "literals_end": int(lits_db),
            "hit_records": int(hits_total), "tuned": tuned, "records": args.records,
            "verified_steps": verified, "steps_in_flight": args.inflight,
            "screen": args.screen,
            "signs": args.signs,
        },
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)),
                       "p99": float(np.percentile(lat_ms, 99)),
                       "p999": float(np.percentile(lat_ms, 99.9)),
                       "max": float(lat_ms.max())},
        "graph_ms": {"mean": float(np.mean(step_ms)), "max": float(np.max(step_ms))},
        "host_submit_ms": {"mean": float(np.mean(submit_s)) * 1e3,
                           "max": float(np.max(submit_s)) * 1e3},
        "step_hit_buffer": e.step_stats(),
    }
    if n_plain:
        plain_rate = sum(plain_checks) / plain_s
        batched_rate = sum(plain_checks) / (done_t[n_plain - 1] - t_start)
        out["plain_form"] = {
            "steps": n_plain, "checks_per_s": plain_rate,
            "batched_checks_per_s_same_steps": batched_rate,
            "batched_over_plain": batched_rate / plain_rate,
            "what": "the first timed steps replayed one at a time through the plain entry "
                    "points (append, apply, sweep, records to the host, roll forward), "
                    "same tuning"}
    print(json.dumps(out), flush=True)
    e.close()
    return 0


def run_serving(args) -> int:
    """serving-32 through the native serving loop (include/gss_serve.h):
    requests published into its ring at their Poisson arrival times, admitted
    on the device by the step that starts next; latency = arrival -> the
    request's records handed to it on the host."""
    from gpusharesat_amd import serving as sv
    if args.records == "clauses":
        log("serving-32 hands each request its records by word: --records full or ids")
        return 2
    cfg = sv.ServeConfig(duration=args.duration, rate=args.rate, max_inflight=args.inflight,
                         records=sv.RECORDS_IDS if args.records == "ids" else sv.RECORDS_FULL,
                         planted=args.signs == "planted", flips=args.flips)
    log(f"serving-32: {cfg}")
    srv = sv.Server(cfg, log=log)
    try:
        if args.tuning:
            srv.srv.set_tuning(args.tuning)
            tuned = args.tuning
        else:
            tuned = None if args.no_autotune else srv.tune()
        preloaded = 0
        if args.preload:
            preloaded = srv.preload(args.preload)
            log(f"preloaded {preloaded} clauses")
        out = srv.run()
    finally:
        srv.close()
    lat = out.latency_ms
    pct = {q: float(np.percentile(lat, q)) for q in (50, 99, 99.9)}
    log(f"{out.requests} requests in {out.steps} steps over {out.elapsed:.1f}s, "
        f"{out.checks_per_s:.3e} checks/s, p50 {pct[50]:.2f} ms p99 {pct[99]:.2f} ms "
        f"p999 {pct[99.9]:.2f} ms, {out.records} records")
    # latency by arrival time (the DB grows through the run)
    windows = []
    for w0 in np.arange(0.0, cfg.duration, 10.0):
        sel = lat[(out.arrive_s >= w0) & (out.arrive_s < w0 + 10.0)]
        if sel.size:
            windows.append({"from_s": float(w0), "requests": int(sel.size),
                            "p50": float(np.percentile(sel, 50)),
                            "p99": float(np.percentile(sel, 99))})
    log("by arrival window: " + ", ".join(
        f"{w['from_s']:.0f}s p50 {w['p50']:.2f} p99 {w['p99']:.2f}" for w in windows))
    signs = ("planted (literals true under a planted assignment w.p. 0.9, assigned values "
             "agree w.p. 0.99)" if cfg.planted else "random (uniform clause signs and "
             "assigned polarities, as config 2)")
    print(json.dumps({
        "metric": METRIC, "value": out.checks_per_s, "unit": "checks/s", "n_gpus": 1,
        "steps": out.steps, "warmup": 0, "ms_per_step": out.elapsed * 1e3 / max(1, out.steps),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"serving-32: 32 solvers over 2M vars, Poisson {cfg.rate:g} "
                               "requests/s of U[64,1024] assignments (30-70% assigned), 200k "
                               "clauses/s of len 2+Geom(0.1) cap 200, 120 GB budget",
                   "signs": signs, "records": args.records,
                   "duration_s": cfg.duration, "requests": out.requests,
                   "served_per_s": out.requests / out.elapsed,
                   "clauses_end": out.clauses_end, "literals_end": out.literals_end,
                   "records_total": out.records, "max_queue": out.max_queue, "tuned": tuned,
                   "steps_in_flight": cfg.max_inflight, "loop": "native (gss_server)",
                   "prefilter": os.environ.get("GPUSHARESAT_SERVE_PREFILTER", "1") != "0",
                   "prefilter_vote": int(os.environ.get("GPUSHARESAT_SERVE_VOTE", "2")),
                   "preloaded_clauses": preloaded, "flips": cfg.flips},
        "latency_ms": {"p50": pct[50], "p99": pct[99], "p999": pct[99.9],
                       "max": float(lat.max()),
                       "what": "arrival -> the request's records handed to it on the host"},
        "latency_by_arrival_10s": windows,
        "device_ms_per_step": out.device_ms / max(1, out.steps),
        "host_grouped_steps": out.host_grouped, "graph_captures": out.graph_captures,
        "step_hit_buffer": out.step_stats,
        "serving_thread": out.thread,
        "driver_lag_ms": {"p99": float(np.percentile(out.lag_ms, 99)),
                          "p999": float(np.percentile(out.lag_ms, 99.9)),
                          "max": float(out.lag_ms.max()),
                          "what": "arrival -> published in the ring by the Python driver "
                                  "(part of latency_ms)"},
        "device_ms_per_step_by_window": [
            {"from_s": round(a[0], 1), "to_s": round(b[0], 1), "clauses_end": int(b[3]),
             "steps": int(b[1] - a[1]),
             "device_ms": (b[2] - a[2]) / (b[1] - a[1]) if b[1] > a[1] else None}
            for a, b in zip(out.samples, out.samples[1:])],
    }), flush=True)
    return 0


def run_perftest(args) -> int:
    """The reference's own benchmark shape (glucose-syrup/perftest/perfTest.cu:
    testPerf) through the drop-in GpuClauseSharer interface: 2M clauses of
    12..20 random literals over 500 variables, one solver; per round the
    solver unsets its previous values, sets each variable true / false /
    unset with probability 1/3 each, sends the assignment, the GPU runs
    (execute = two gpuRun calls, as the reference's test helper) and the
    solver imports its reported clauses.  Seeded numpy instance of that
    shape (the exact golden-count instance is checked in the tests).
    Reports the reference's two figures: executions per second (rounds over
    GPU + import time) and reads per microsecond (rounds x clauses x (1 +
    mean length) / GPU time)."""
    import ctypes
    from gpusharesat_amd import _lib
    from gpusharesat_amd.sharer import GpuClauseSharer, default_options
    rounds, warm = args.steps, args.warmup
    V, n_cl, lo, hi = 500, 2_000_000, 12, 20
    rng = np.random.default_rng(25)
    lens = rng.integers(lo, hi + 1, size=n_cl).astype(np.int32)
    total = int(lens.sum())
    vars_ = rng.integers(0, V, size=total)
    lits = (2 * vars_ + rng.integers(0, 2, size=total)).astype(np.int32)  # mkLit(var, sign)
    off = np.zeros(n_cl + 1, np.int64)
    off[1:] = np.cumsum(lens)
    log(f"perftest: {n_cl} clauses ({total} literals) over {V} vars, {rounds} rounds")
    s = GpuClauseSharer(default_options(min_gpu_latency_micros=0, verbosity=0))
    try:
        s.setVarCount(V)
        s.setCpuSolverCount(1)
        t0 = time.perf_counter()
        base = lits.ctypes.data
        cid = ctypes.c_int64()
        add = _lib.load().gss_sharer_add_clause
        for i in range(n_cl):
            rc = add(s._h, -1, base + 4 * int(off[i]), int(lens[i]), ctypes.byref(cid))
            if rc:
                raise RuntimeError(f"addClause {i} failed: {_lib.load().gss_last_error()}")
        log(f"added in {time.perf_counter() - t0:.1f}s")
        prev = np.zeros(0, np.int32)
        gpu_s = imp_s = 0.0
        reports = 0
        for r in range(warm + rounds):
            p = rng.integers(0, 3, size=V)
            vs = np.nonzero(p < 2)[0]
            cur = (2 * vs + (p[vs] == 1)).astype(np.int32)  # p 0: true, 1: false
            if prev.size:
                s.unsetSolverValues(0, prev)
            if not s.trySetSolverValues(0, cur):
                raise RuntimeError("trySetSolverValues refused")
            s.trySendAssignment(0)
            prev = cur
            t_a = time.perf_counter()
            s.execute()
            t_b = time.perf_counter()
            got = s.popAll(0)
            t_c = time.perf_counter()
            if r >= warm:
                gpu_s += t_b - t_a
                imp_s += t_c - t_b
                reports += len(got)
        execs = rounds / (gpu_s + imp_s)
        reads_us = rounds * n_cl * (1 + (lo + hi) / 2) / (gpu_s * 1e6)
        log(f"{rounds} rounds: gpu {gpu_s * 1e3:.0f} ms, import {imp_s * 1e3:.0f} ms, "
            f"{reports} reports; executions per second {execs:.0f}, reads per microsecond "
            f"{reads_us:.0f}")
    finally:
        s.close()
    print(json.dumps({
        "metric": "perfTest executions per second (GpuClauseSharer: set, send, gpuRun x2, import)",
        "value": execs, "unit": "rounds/s", "n_gpus": 1, "steps": rounds, "warmup": warm,
        "ms_per_step": (gpu_s + imp_s) * 1e3 / rounds, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "perftest: 2M clauses len U[12,20] over 500 vars, 1 solver, "
                               "1 assignment per round (1/3 true, 1/3 false, 1/3 unset)",
                   "reports": reports},
        "reads_per_microsecond": reads_us,
    }), flush=True)
    return 0


def run_ceilings(args) -> int:
    """Device properties, a plain copy, and random 128-byte line gathers from
    tables of several sizes (the 16-word sweep's access pattern): what the
    device reaches, to set the sweep's measured traffic against."""
    import ctypes
    from gpusharesat_amd import _lib

    class DeviceInfo(ctypes.Structure):
        _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 64),
                    ("cus", ctypes.c_int32), ("clock_khz", ctypes.c_int32),
                    ("l2_bytes", ctypes.c_int64), ("lds_per_cu", ctypes.c_int64),
                    ("hbm_bytes", ctypes.c_int64), ("mem_clock_khz", ctypes.c_int32),
                    ("mem_bus_bits", ctypes.c_int32)]

    info = DeviceInfo()
    _lib.call("gss_device_query", 0, ctypes.byref(info))
    dev = {"name": info.name.decode(), "arch": info.arch.decode(), "cus": info.cus,
           "clock_mhz": info.clock_khz / 1e3, "l2_bytes": info.l2_bytes,
           "lds_per_cu": info.lds_per_cu, "hbm_bytes": info.hbm_bytes,
           "mem_clock_mhz": info.mem_clock_khz / 1e3, "mem_bus_bits": info.mem_bus_bits}
    log(f"device: {dev}")
    # PCIe first: after large device allocations have come and gone, the same
    # copies have been seen at about half the rate; the best of two runs
    d2h, h2d = ctypes.c_double(), ctypes.c_double()
    pcie_bytes = 424 << 20  # the records of one sweep-1024
    best = [0.0, 0.0]
    for _ in range(2):
        _lib.call("gss_ceiling_pcie", 0, pcie_bytes, 5, ctypes.byref(d2h), ctypes.byref(h2d))
        best = [max(best[0], d2h.value), max(best[1], h2d.value)]
    pcie = {"bytes": pcie_bytes, "d2h_gbps": best[0], "h2d_gbps": best[1]}
    log(f"pcie {pcie_bytes >> 20} MiB to pinned memory: {best[0]:.1f} GB/s device->host, "
        f"{best[1]:.1f} GB/s host->device")
    gbps, ms = ctypes.c_double(), ctypes.c_double()
    copy_bytes = 4 << 30
    _lib.call("gss_ceiling_copy", 0, copy_bytes, 10, ctypes.byref(gbps), ctypes.byref(ms))
    copy = {"bytes": copy_bytes, "gbps": gbps.value, "frac": gbps.value / HBM_PEAK_GBS,
            "ms": ms.value}
    log(f"copy {copy_bytes >> 20} MiB: {gbps.value:.0f} GB/s read+write "
        f"({gbps.value / HBM_PEAK_GBS:.2f} of peak)")
    gathers = []
    reads = 32 << 20  # 4 GiB of lines per run
    for table in (2 << 20, 32 << 20, 256 << 20, 1 << 30, 8 << 30):
        for unroll in (4, 8):
            _lib.call("gss_ceiling_gather", 0, table, reads, unroll, 5, ctypes.byref(gbps),
                      ctypes.byref(ms))
            gathers.append({"table_bytes": table, "unroll": unroll, "gbps": gbps.value,
                            "frac": gbps.value / HBM_PEAK_GBS, "ms": ms.value})
            log(f"gather {table >> 20:5d} MiB table, {unroll} lines in flight per lane: "
                f"{gbps.value:.0f} GB/s of lines")
    # reads of part of a line, or of two lines, from the 256 MiB table: what a
    # half-row (512 slots) or a row of both signs (2048 slots) would cost
    blocks = []
    for block in (64, 128, 256):
        _lib.call("gss_ceiling_gather_block", 0, 256 << 20, block, reads, 4, 5,
                  ctypes.byref(gbps), ctypes.byref(ms))
        blocks.append({"table_bytes": 256 << 20, "block_bytes": block, "unroll": 4,
                       "gbps": gbps.value, "reads_per_s": reads / (ms.value * 1e-3),
                       "ms": ms.value})
        log(f"gather 256 MiB table, random {block}-byte blocks: {gbps.value:.0f} GB/s, "
            f"{reads / (ms.value * 1e-3) / 1e9:.1f} G reads/s")
    out = {"metric": "device memory ceilings (copy; random 128-byte line gathers)",
           "value": copy["gbps"], "unit": "GB/s", "n_gpus": 1, "higher_is_better": True,
           "data": "synthetic", "config": {"workload": "ceilings", "device": dev},
           "copy": copy, "pcie": pcie, "gather": gathers, "gather_blocks": blocks}
    print(json.dumps(out), flush=True)
    return 0


def run_residency(args) -> int:
    from gpusharesat_amd import residency as rs
    s = args.res_scale
    cfg = rs.SimConfig(var_count=int(2_000_000 * s), arrivals=int(60_000_000 * s),
                       budget=int(20_000_000 * s))
    log(f"residency: {cfg}")
    res = {}
    for pol, name in ((rs.POLICY_ACTIVITY, "activity"), (rs.POLICY_LEARNED, "learned")):
        t0 = time.time()
        out = rs.simulate(cfg, pol, log=log)
        res[name] = out
        log(f"{name}: {out.hits_per_gb:.4e} hit pairs/GB, {out.reductions} reductions, "
            f"{time.time() - t0:.1f}s")
    a, l = res["activity"], res["learned"]
    print(json.dumps({
        "metric": "hit pairs per GB of clause literals checked (residency policy)",
        "value": l.hits_per_gb, "unit": "hit pairs/GB", "n_gpus": 1, "steps": cfg.rounds,
        "warmup": 0, "higher_is_better": True, "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "residency: 60M arrivals over 2M vars into a 20M-clause budget, "
                               "16 solvers x 64 slots, hot sets of 5k vars redrawn every 50 "
                               "rounds", "scale": s},
        "activity": {"hits_per_gb": a.hits_per_gb, "hit_pairs": a.hit_pairs,
                     "bytes_checked": a.bytes_checked, "reductions": a.reductions},
        "learned": {"hits_per_gb": l.hits_per_gb, "hit_pairs": l.hit_pairs,
                    "bytes_checked": l.bytes_checked, "reductions": l.reductions,
                    "train_loss": l.train_loss},
        "gain": l.hits_per_gb / a.hits_per_gb if a.hits_per_gb else None,
    }), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
