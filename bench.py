#!/usr/bin/env python3
"""Benchmark: TV refinement throughput on one MI355X.

Workload (one "step" = one refinement pass over the whole hypergraph):
  column-net hypergraph of a 1,000,000 x 1,000,000 sparse matrix, total row
  degree ~ Zipf(alpha=2.1) on [2, 5000] with the diagonal always present (net =
  row, source = diagonal vertex, vertex = column), unit weights, K = 256,
  eps = 0.03, balanced deterministic round schedule; initial partition = the
  streaming LDG (host). Inputs are resident in HBM before the timed region.

Metric: pin-visits/s = steps * pins / seconds (a pin-visit is one (vertex,
incident net) incidence scored in a pass).

Also reported on the same JSON line:
  roofline      the evaluation kernel (k_tv_eval), timed live with HIP events
                on the engine's stream, against the 8 TB/s HBM peak
  cpu_baseline  the reference's own OpenMP refinement (partition_par,
                TV/partition.cpp:512-981, built from its sources in oracle/_ref)
                on the host cores, a bounded sample of the same workload
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


def lambda_visits(hg, lam) -> int:
    sizes = np.diff(hg.net_ptr.astype(np.int64))
    return int(np.dot(sizes, lam.astype(np.int64)))


def cpu_info():
    """CPU model, the machine's physical cores and the threads this process
    may run on."""
    model, cores = "unknown", set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for ln in f:
                k, _, v = ln.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    return {"model": model, "physical_cores": len(cores) or None, "affinity": len(os.sched_getaffinity(0))}


# The box's CPU share for one GPU: the reference's OpenMP baseline runs on at
# most this many threads (all of them, bound to cores).
CPU_SHARE = 16


def _pmc_values(csv_dir, counter, kernel):
    """Per-dispatch values of a PMC counter for the dispatches of `kernel`
    (rocprofv3 counter_collection.csv)."""
    import csv
    import glob

    vals = {}
    for path in glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") == counter and kernel in row.get("Kernel_Name", ""):
                    key = row.get("Dispatch_Id")
                    vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
    return list(vals.values())


# FETCH_SIZE / WRITE_SIZE read, for a plain device-to-device copy of known
# size, half / all of the bytes the copy moves (three 1 GiB copies: 1,572,909
# KiB fetched, 3,145,728 KiB written; calibration run and command in
# profiles/tv_eval_c2_1M_K256_R262144.md): the factors that turn the counters
# into HBM bytes.
COUNTER_FACTOR = {"FETCH_SIZE": 2.0, "WRITE_SIZE": 1.0}


# the LDS and atomic counters reported beside the traffic, one rocprofv3 pass:
# LDS instructions, LDS atomic instructions, LDS bank-conflict cycles, and the
# atomic requests that reach the L2 (global atomics)
LDS_ATOMIC_COUNTERS = ("SQ_INSTS_LDS", "SQ_INSTS_LDS_ATOMIC", "SQ_LDS_BANK_CONFLICT", "TCC_ATOMIC_sum")


def measure_traffic(args):
    """HBM bytes per k_tv_eval launch, measured in this run: rocprofv3 --pmc
    FETCH_SIZE, then WRITE_SIZE (KiB per dispatch) over a 2-pass probe of the
    same workload (--traffic-probe), each scaled by COUNTER_FACTOR; then the
    LDS and atomic counters per dispatch in a third pass. Returns (bytes,
    detail)."""
    import glob
    import shutil
    import subprocess
    import tempfile

    rocprof = shutil.which("rocprofv3")
    if not rocprof:
        return None, {"error": "rocprofv3 not found"}
    env = dict(os.environ, TMPDIR="/tmp")
    out = {}
    try:
        with tempfile.TemporaryDirectory(dir="/tmp") as d:
            for counter in ("FETCH_SIZE", "WRITE_SIZE"):
                odir = os.path.join(d, counter)
                cmd = [rocprof, "--pmc", counter, "-f", "csv", "-d", odir, "-o", "run", "--", sys.executable, "-u",
                       os.path.abspath(__file__), "--traffic-probe", "--round", str(args.round), "--sched",
                       args.sched]
                p = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd="/tmp", env=env)
                if not glob.glob(os.path.join(odir, "**", "*counter_collection.csv"), recursive=True):
                    return None, {"error": f"rocprofv3 --pmc {counter} rc={p.returncode}: {p.stderr[-300:]}"}
                kv = _pmc_values(odir, counter, "k_tv_eval")
                if not kv:
                    return None, {"error": f"no {counter} values for k_tv_eval"}
                out[counter] = {"kib_per_dispatch": round(sum(kv) / len(kv), 1), "dispatches": len(kv),
                                "factor": COUNTER_FACTOR[counter]}
            odir = os.path.join(d, "lds_atomic")
            cmd = [rocprof, "--pmc", *LDS_ATOMIC_COUNTERS, "-f", "csv", "-d", odir, "-o", "run", "--",
                   sys.executable, "-u", os.path.abspath(__file__), "--traffic-probe", "--round", str(args.round),
                   "--sched", args.sched]
            p = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd="/tmp", env=env)
            la = {}
            for counter in LDS_ATOMIC_COUNTERS:
                kv = _pmc_values(odir, counter, "k_tv_eval")
                if kv:
                    la[counter] = round(sum(kv) / len(kv), 1)
            traffic = sum(v["kib_per_dispatch"] * 1024.0 * v["factor"] for v in out.values())
            out["lds_atomic_per_dispatch"] = la if la else {"error": f"rocprofv3 rc={p.returncode}"}
    except Exception as e:  # reported, never fatal for the GPU number
        return None, {"error": str(e)[:300]}
    return int(traffic), out


def cpu_baseline(hg, K, passes, threads):
    from oracle import oracle as O

    if not O.ref_available("tv"):
        return None
    # partition_par reads its per-thread move counts before every thread has
    # written them (TV/partition.cpp:806-814; the array is uninitialized,
    # :671), which can crash the reference; a crashed run is retried, then
    # the thread count is halved.
    crashes = []
    t = threads
    while t >= 1:
        for _ in range(2):
            env = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores", "OMP_NUM_THREADS": str(t)}
            try:
                r = O.run_reference("tv", hg, K, passes, t, timeout=900, env_extra=env)
            except RuntimeError as e:
                crashes.append(f"{t} threads: " + ("segfault" if "rc=-11" in str(e) else str(e)[-40:]))
                continue
            secs = r["refine_seconds"]
            after = [ln for ln in r["stdout"].splitlines() if ln.startswith("After Refinement TV")]
            note = f"; crashed runs before this one: {crashes}" if crashes else ""
            return {
                "value": passes * hg.pins / secs,
                "unit": "pin-visits/s",
                "cores": t,
                "kind": "reference",
                "sample": (f"reference partition_par (TV/partition.cpp:512-981, built from its sources) on the "
                           f"same 1M-vertex instance, K={K}, its hard-coded eps=0.05 constants, {passes} passes, "
                           f"{t} OpenMP threads; 'Refinement took' = {secs:.3f} s; "
                           f"{after[-1] if after else ''}{note}"),
            }
        t //= 2
    return {"error": f"reference crashed at every thread count: {crashes}"}


OBJECTIVES = ("TV", "CN", "TM", "MSM", "MSV", "MSRV")


def batch_specs(count, seed):
    """Instances of the batched workload: vertex count log-uniform in
    [1,000, 200,000], K in {16, 64, 256}, objective uniform over the six."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.floor(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))))
        out.append((i, n, int(rng.choice([16, 64, 256])), OBJECTIVES[int(rng.integers(6))]))
    return out


def batch_schedule(metric, nov, K, C, max_pw, min_pw, cap):
    """Round schedule per objective: TV the balanced schedule, TM a whole pass
    per round (its reference schedule), the max-set objectives the reference
    round schedule with rounds of 16384 vertices."""
    from fastpartitioner_amd.engine import SCHED_BAL, SCHED_REF, schedule

    if metric == "CN":
        return None
    if metric == "TV":
        return schedule(SCHED_BAL, 16384, C, cap, 0, cap)
    R = nov if metric == "TM" else 16384
    return schedule(SCHED_REF, R, C, max_pw, min_pw)


def run_batch(args):
    """Batched refinement of many hypergraphs against one-at-a-time refinement
    of the same hypergraphs from the same states: hypergraphs/s of each, and
    whether every final part vector is identical."""
    from concurrent.futures import ThreadPoolExecutor

    # hardware queues of this process, read when the HIP runtime starts (more
    # than the runtime's default of 4 made the batch and the one-at-a-time
    # loop slower alike on MI355X)
    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    from fastpartitioner_amd import hypergraph as H
    from fastpartitioner_amd.engine import Batch, Engine

    eps = 0.05
    specs = batch_specs(args.instances, args.seed)

    def prep(spec):
        # each member starts from its objective's own initial stream
        i, n, K, metric = spec
        hg = H.colnet(n, H.DEG_UNIFORM, 4, 32, seed=1000 + i)
        C, max_pw, min_pw = H.ref_constants(hg.nov, K, eps)
        part = H.initial_partition_for(metric, hg, K, eps)
        return hg, part, batch_schedule(metric, hg.nov, K, C, max_pw, min_pw, H.balance_cap(hg.nov, K, eps))

    t = time.time()
    with ThreadPoolExecutor(max_workers=min(16, len(os.sched_getaffinity(0)))) as ex:
        prepared = list(ex.map(prep, specs))
    pins = sum(hg.pins for hg, _, _ in prepared)
    print(f"# batch: {len(specs)} instances, {sum(hg.nov for hg, _, _ in prepared)} vertices, {pins} pins, "
          f"prep {time.time() - t:.1f}s", file=sys.stderr, flush=True)

    engines = []
    try:
        t = time.time()
        for (i, n, K, metric), (hg, part, sched) in zip(specs, prepared):
            eng = Engine(0)
            engines.append(eng)
            eng.upload(hg)
            eng.set_partition(K, part)
        print(f"# batch: uploaded in {time.time() - t:.1f}s", file=sys.stderr, flush=True)
        # warmup passes, untimed on both sides: the batch spends them on the
        # autotuner trials and the graph capture, so the timed passes measure
        # the steady state of each
        wb = max(args.warmup, args.tune_passes + 1)
        # one at a time
        for eng, (i, n, K, metric), (hg, part, sched) in zip(engines, specs, prepared):
            eng.refine(metric, wb, sched)
        t0 = time.perf_counter()
        for eng, (i, n, K, metric), (hg, part, sched) in zip(engines, specs, prepared):
            eng.refine(metric, args.passes, sched)
        t_solo = time.perf_counter() - t0
        solo_parts = [eng.get_state().part for eng in engines]
        print(f"# batch: one at a time {t_solo:.2f}s", file=sys.stderr, flush=True)
        for eng, (i, n, K, metric), (hg, part, sched) in zip(engines, specs, prepared):
            eng.set_partition(K, part)
        # batched
        with Batch(0) as b:
            b.set_lanes(args.lanes)
            for eng, (i, n, K, metric), (hg, part, sched) in zip(engines, specs, prepared):
                b.add(eng, metric, sched)
            t0 = time.perf_counter()
            b.refine(wb, tune_passes=args.tune_passes)
            t_first = time.perf_counter() - t0
            t0 = time.perf_counter()
            dev_ms = b.refine(args.passes)
            t_batch = time.perf_counter() - t0
            shapes = {}
            for j in range(len(engines)):
                mm = b.member(j)
                key = f"{mm['eval_waves']}x{mm['eval_grid']}"
                shapes[key] = shapes.get(key, 0) + 1
        print(f"# batch: batched {t_batch:.2f}s (warmup with tuning and capture {t_first:.2f}s)", file=sys.stderr,
              flush=True)
        identical = all(np.array_equal(eng.get_state().part, sp) for eng, sp in zip(engines, solo_parts))
    finally:
        for eng in engines:
            eng.close()
    n = len(specs)
    out = {
        "metric": "batched refinement hypergraphs/s (mixed size, K and objective)",
        "value": round(n / t_batch, 2),
        "unit": "hypergraphs/s",
        "n_gpus": 1,
        "steps": args.passes,
        "warmup": wb,
        "ms_per_step": round(t_batch * 1e3 / args.passes, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"batch-{n}-colnet-uniform4to32-logU1e3to2e5", "instances": n, "pins": int(pins),
                   "passes": args.passes, "tune_passes": args.tune_passes, "lanes": args.lanes,
                   "hw_queues": args.hw_queues,
                   "objectives": {m: sum(1 for s in specs if s[3] == m) for m in OBJECTIVES},
                   "tuned_shapes": shapes},
        "one_at_a_time_hgs": round(n / t_solo, 2),
        "speedup_over_one_at_a_time": round(t_solo / t_batch, 2),
        "identical_results": bool(identical),
        "batch_device_ms": round(dev_ms, 2),
        "warmup_s_with_tuning_and_capture": round(t_first, 3),
    }
    print(json.dumps(out), flush=True)


def run_serve(args):
    """The serving loop under Poisson arrivals: resident hypergraphs, requests
    with mixed K and objective admitted by HBM footprint; p50/p99 latency
    (arrival to completion) and completed hypergraphs/s. A sample of requests
    is rerun alone to check that results do not depend on the mix."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    from fastpartitioner_amd import hypergraph as H
    from fastpartitioner_amd.engine import Engine, Server

    eps, Ks = 0.05, (32, 128, 512)
    rng = np.random.default_rng(args.seed)
    sizes = [int(np.floor(np.exp(rng.uniform(np.log(5e3), np.log(1e6))))) for _ in range(args.graphs)]

    def prep_graph(t):
        return H.colnet(sizes[t], H.DEG_ZIPF, 2, 5000, 2.3, seed=2000 + t)

    t = time.time()
    workers = min(16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        graphs = list(ex.map(prep_graph, range(args.graphs)))

    # a request starts from its objective's own initial stream (TV LDG2, the
    # CN stream, band-free LDG for MSV/MSRV, the message greedy for TM/MSM)
    def prep_part(key):
        g, K, mode = key
        hg = graphs[g]
        C, max_pw, min_pw = H.ref_constants(hg.nov, K, eps)
        return H.initial_partition_mode(hg, mode, K, C, max_pw, min_pw)

    modes = sorted(set(H.INIT_OF.values()))
    keys = [(g, K, m) for g in range(args.graphs) for K in Ks for m in modes]
    with ThreadPoolExecutor(max_workers=workers) as ex:
        parts = dict(zip(keys, ex.map(prep_part, keys)))

    def part_of(g, K, metric):
        return parts[(g, K, H.INIT_OF[metric])]
    # arrivals: Poisson process at args.rate requests/s for args.duration s
    arrivals = []
    at = rng.exponential(1.0 / args.rate)
    while at < args.duration:
        g = int(rng.integers(args.graphs))
        K = int(rng.choice(Ks))
        metric = OBJECTIVES[int(rng.integers(6))]
        arrivals.append((at, g, K, metric))
        at += rng.exponential(1.0 / args.rate)
    pins = [hg.pins for hg in graphs]
    print(f"# serve: {args.graphs} hypergraphs, {sum(hg.nov for hg in graphs)} vertices, {sum(pins)} pins, "
          f"{len(arrivals)} requests over {args.duration}s, prep {time.time() - t:.1f}s", file=sys.stderr, flush=True)

    def sched_of(g, K, metric):
        hg = graphs[g]
        C, max_pw, min_pw = H.ref_constants(hg.nov, K, eps)
        return batch_schedule(metric, hg.nov, K, C, max_pw, min_pw, H.balance_cap(hg.nov, K, eps))

    with Server(0, hbm_budget=args.budget_gib << 30, workers=args.workers) as srv:
        srv.set_replicas(args.replicas)
        for hg in graphs:
            srv.add_graph(hg)
        rids = [None] * len(arrivals)
        scheds = [sched_of(g, K, metric) for _, g, K, metric in arrivals]
        start = time.perf_counter()

        def submitter():
            last = 0.0
            for i, (a, g, K, metric) in enumerate(arrivals):
                delay = start + a - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                rids[i] = srv.submit(g, metric, K, part_of(g, K, metric), scheds[i], args.passes)
                if a - last >= 10:
                    last = a
                    print(f"# serve: {i + 1} submitted at {a:.0f}s", file=sys.stderr, flush=True)

        th = threading.Thread(target=submitter)
        th.start()
        infos = []
        for i in range(len(arrivals)):
            while rids[i] is None:
                time.sleep(0.001)
            info, part = srv.wait(rids[i], with_part=i < args.check)
            infos.append((info, part))
        th.join()
        stats = srv.stats()
    lat = np.array([info["end_ms"] - info["submit_ms"] for info, _ in infos])
    first = min(info["submit_ms"] for info, _ in infos)
    last_end = max(info["end_ms"] for info, _ in infos)
    span_s = (last_end - first) / 1e3
    # the first requests rerun alone on fresh contexts
    same = True
    for (info, part), (a, g, K, metric) in zip(infos[:args.check], arrivals[:args.check]):
        with Engine(0) as e:
            e.upload(graphs[g])
            e.set_partition(K, part_of(g, K, metric))
            e.refine(metric, args.passes, sched_of(g, K, metric))
            same = same and np.array_equal(e.get_state().part, part) and e.metrics() == info["metrics"]
    out = {
        "metric": "serving hypergraphs/s and latency (Poisson arrivals, mixed K and objective)",
        "value": round(len(infos) / span_s, 2),
        "unit": "hypergraphs/s",
        "n_gpus": 1,
        "steps": args.passes,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"serve-{args.graphs}-colnet-zipf2.3-logU5e3to1e6-poisson{args.rate}",
                   "requests": len(infos), "rate_per_s": args.rate, "duration_s": args.duration,
                   "passes": args.passes, "K": list(Ks), "workers": args.workers, "replicas": args.replicas,
                   "budget_gib": args.budget_gib,
                   "hw_queues": args.hw_queues, "resident_pins": int(sum(pins))},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2),
                       "mean": round(float(lat.mean()), 2), "max": round(float(lat.max()), 2)},
        "peak_inflight": stats["peak_inflight"],
        "peak_inflight_gib": round(stats["peak_bytes"] / 2**30, 3),
        "same_alone_and_in_mix": bool(same),
        "checked_requests": args.check,
    }
    print(json.dumps(out), flush=True)


def tv1m_instance(args):
    from fastpartitioner_amd import hypergraph as H
    from fastpartitioner_amd.engine import SCHED_BAL, SCHED_REF, schedule

    K, eps = 256, 0.03
    hg = H.workload("c2", seed=1, scale=args.scale)
    W = hg.total_weight()
    C, max_pw, min_pw = H.ref_constants(W, K, eps)
    cap = H.balance_cap(W, K, eps)
    t = time.perf_counter()
    part0 = H.initial_partition(hg, K, C, max_pw, min_pw)
    args.host_stream_s = time.perf_counter() - t
    args.ldg = (C, max_pw)
    if args.sched == "bal":
        sched = schedule(SCHED_BAL, args.round, C, cap, 0, cap)
    else:
        sched = schedule(SCHED_REF, args.round, C, max_pw, min_pw)
    return hg, K, eps, cap, part0, sched


def cn_pass_bytes(hg, part, lam):
    """Algorithmic bytes of one CN pass's evaluation on a state (the per-unit
    model of the lambda = 2 push, CN/partition.cpp:629-690): 6 B per net (its
    offset and lambda), 8 B per pin of a lambda = 2 net (pin id and part), and
    6 B per incident net of each pin on the smaller side (net id and lambda)
    -- the uncut-net check."""
    sizes = np.diff(hg.net_ptr.astype(np.int64))
    two = np.nonzero(lam == 2)[0]
    if len(two) == 0:
        return 6 * hg.noe
    starts, lens = hg.net_ptr[two].astype(np.int64), sizes[two]
    idx = np.repeat(starts, lens) + (np.arange(int(lens.sum())) - np.repeat(np.cumsum(lens) - lens, lens))
    pins = hg.net_pins[idx]
    q = part[pins]
    first = np.repeat(q[np.cumsum(lens) - lens], lens)
    is_a = q == first
    ca = np.add.reduceat(is_a.astype(np.int64), np.cumsum(lens) - lens)
    small_is_a = np.repeat(ca < lens - ca, lens)
    in_small = is_a == small_is_a
    deg = np.diff(hg.vtx_ptr.astype(np.int64))
    return int(6 * hg.noe + 8 * lens.sum() + 6 * deg[pins[in_small]].sum())


def run_tv1m(args):
    from fastpartitioner_amd.engine import Engine, pinned_copy

    t = time.time()
    hg, K, eps, cap, part0, sched = tv1m_instance(args)
    print(f"# workload: nov={hg.nov} pins={hg.pins} K={K} prep {time.time() - t:.1f}s", file=sys.stderr, flush=True)

    if args.traffic_probe:
        # the 2-pass probe that rocprofv3 --pmc runs (measure_traffic)
        with Engine(0) as eng:
            eng.upload(hg)
            eng.set_partition(K, part0)
            eng.refine("TV", 2, sched)
        print(json.dumps({"traffic_probe": True}))
        return

    with Engine(0) as eng:
        # host -> device: both CSR views and the state of the initial
        # partition (pageable host buffers), timed for the PCIe-inclusive rate
        t = time.perf_counter()
        eng.upload(hg)
        eng.set_partition(K, part0)
        t_up = time.perf_counter() - t
        if args.warmup:
            eng.refine("TV", args.warmup, sched)
            eng.set_partition(K, part0)
        m0 = eng.metrics(vectors=True)
        lam0 = lambda_visits(hg, m0["lam"])
        eng.set_kernel_timing(True)
        t0 = time.perf_counter()
        st = eng.refine("TV", args.steps, sched)
        t1 = time.perf_counter()
        launches, eval_ms = eng.kernel_timing()
        eng.set_kernel_timing(False)
        m1 = eng.metrics(vectors=True)
        lam1 = lambda_visits(hg, m1["lam"])
        t = time.perf_counter()
        sizes = eng.get_state().sizes
        t_down = time.perf_counter() - t
        # the same instance's CN refinement (the lambda = 2 push) from the same
        # initial partition, as many passes, timed the same way
        eng.set_partition(K, part0)
        mc0 = eng.metrics(vectors=True)
        cn0 = mc0["CN"]
        if args.warmup:
            eng.refine("CN", args.warmup)
            eng.set_partition(K, part0)
        tc0 = time.perf_counter()
        cst = eng.refine("CN", args.steps)
        tc1 = time.perf_counter()
        cn1 = eng.metrics()["CN"]
        # the evaluation kernel alone, timed over a few evaluations of the final state
        te0 = time.perf_counter()
        for _ in range(10):
            eng.metrics()
        te1 = time.perf_counter()
        # both sides of the refinement on the device: the BFS ordering (from
        # glibc's first rand() % nov, as TV/tv.cpp:120) and the streaming
        # initial partition (LDG2, rounds of 16384) from which TV refines
        # each twice: the first call in the process also pays the first
        # launch of each of its kernels (reported as cold), the second is a
        # repeated use
        bst, t_bfs = [], []
        for _ in range(2):
            t = time.perf_counter()
            bst.append(eng.bfs_ordering(1804289383 % hg.nov)[3])
            t_bfs.append(time.perf_counter() - t)
        C, max_pw = args.ldg
        ist, t_init = [], []
        for _ in range(2):
            t = time.perf_counter()
            ist.append(eng.initial_partition(0, K, 16384, C, max_pw, cap))
            t_init.append(time.perf_counter() - t)
        tv_dev_init = eng.metrics()["TV"]
    # the upload again into fresh contexts (the process's first-use costs
    # paid): from pageable arrays through the staging buffers, and from a
    # reader that filled both CSR views in pinned memory (direct DMA)
    hp = pinned_copy(hg)
    t_up2 = {}
    for name, h in (("pageable", hg), ("pinned", hp)):
        with Engine(0) as eng2:
            t = time.perf_counter()
            eng2.upload(h)
            eng2.set_partition(K, part0)
            t_up2[name] = time.perf_counter() - t
    del hp
    secs = t1 - t0
    value = args.steps * hg.pins / secs

    # algorithmic bytes per pass of the evaluation step, per-unit model of the
    # refinement's data in the reference layout (int32 part + int32 count per
    # list entry), each byte counted once per use, no cache credit:
    # 12 B per vertex, 12 B per (vertex, net) incidence, 8 B per list entry
    # visited (Lambda = sum over nets of |n| * lambda(n))
    lam_avg = 0.5 * (lam0 + lam1)
    bytes_pass = 12 * hg.nov + 12 * hg.pins + 8 * lam_avg
    avg_launch_ms = eval_ms / max(launches, 1)
    bytes_launch = bytes_pass * args.steps / max(launches, 1)
    achieved = bytes_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic, traffic_detail = (None, {"skipped": "--no-traffic"}) if args.no_traffic else measure_traffic(args)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_tv_eval", "launches": launches, "avg_launch_us": round(avg_launch_ms * 1e3, 2),
            "bytes_per_launch": int(bytes_launch), "traffic_counters": traffic_detail,
            "pass_frac": round(bytes_pass / (secs / args.steps) / 1e9 / HBM_PEAK_GBS, 4)}

    cpu = None
    info = cpu_info()
    if not args.no_cpu:
        threads = args.cpu_threads or min(CPU_SHARE, info["affinity"], info["physical_cores"] or CPU_SHARE)
        try:
            cpu = cpu_baseline(hg, K, args.cpu_passes, threads)
            if cpu and "value" in cpu:
                cpu["cpu_model"] = info["model"]
                cpu["machine_physical_cores"] = info["physical_cores"]
                cpu["passes_measured"] = args.cpu_passes
                cpu["extrapolated_20_pass_s"] = round(20 * hg.pins / cpu["value"], 2)
                one = cpu_baseline(hg, K, 1, 1)
                if one and "value" in one:
                    cpu["one_thread"] = {"value": one["value"], "sample": one["sample"]}
        except Exception as e:  # reported, never fatal for the GPU number
            cpu = {"error": str(e)[:300]}

    cn_bytes = cn_pass_bytes(hg, part0, mc0["lam"])
    cn_ms = (tc1 - tc0) * 1e3 / args.steps
    out = {
        "metric": "refine pin-visits/s (1M-vtx, K=256, TV)",
        "value": round(value, 1),
        "unit": "pin-visits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "colnet-zipf2.1-1M-K256-TV", "nov": hg.nov, "pins": hg.pins, "K": K,
                   "eps": eps, "schedule": args.sched, "round": args.round, "passes": args.steps,
                   "lambda_visits_per_pass": int(lam_avg),
                   "tv_initial": int(m0["TV"]), "tv_refined": int(m1["TV"]),
                   "max_part": int(sizes.max()), "cap": int(cap),
                   "moves": int(st.moves), "rounds": int(st.rounds),
                   "device_ms": round(st.device_ms, 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "cn_same_instance": {"passes": args.steps, "ms_per_step": round(cn_ms, 4),
                             "pin_visits_per_s": round(args.steps * hg.pins / (tc1 - tc0), 1),
                             "cn_initial": int(cn0), "cn_refined": int(cn1), "moves": int(cst.moves),
                             "roofline": {"bound": "hbm", "scope": "whole pass", "bytes_per_pass": cn_bytes,
                                          "achieved": round(cn_bytes / (cn_ms * 1e-3) / 1e9, 1),
                                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                          "frac": round(cn_bytes / (cn_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
        "metrics_evaluation": {"ms": round((te1 - te0) * 1e3 / 10, 4),
                               "bytes": int(8 * hg.pins + 12 * hg.noe)},
        "pcie_inclusive": {"upload_s": round(t_up, 4), "state_download_s": round(t_down, 4),
                           "pin_visits_per_s": round(args.steps * hg.pins / (secs + t_up + t_down), 1),
                           "upload_again_pageable_s": round(t_up2["pageable"], 4),
                           "upload_again_pinned_s": round(t_up2["pinned"], 4),
                           "pin_visits_per_s_pinned": round(
                               args.steps * hg.pins / (secs + t_up2["pinned"] + t_down), 1)},
        "host_path_on_device": {
            "bfs_ordering": {"device_ms": round(bst[1].device_ms, 3), "wall_ms": round(t_bfs[1] * 1e3, 3),
                             "cold_device_ms": round(bst[0].device_ms, 3),
                             "cold_wall_ms": round(t_bfs[0] * 1e3, 3),
                             "levels": int(bst[1].levels), "reached": int(bst[1].reached)},
            "initial_partition": {"stream": "LDG2 with bands", "round": 16384,
                                  "device_ms": round(ist[1].device_ms, 3), "wall_ms": round(t_init[1] * 1e3, 3),
                                  "cold_device_ms": round(ist[0].device_ms, 3),
                                  "cold_wall_ms": round(t_init[0] * 1e3, 3),
                                  "subrounds": int(ist[1].subrounds), "tv": int(tv_dev_init),
                                  "tv_sequential_stream": int(m0["TV"]),
                                  "host_sequential_stream_s": round(args.host_stream_s, 3)}},
    }
    if cpu and "value" in cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out), flush=True)


C3_OBJECTIVES = ("TM", "MSM", "MSV", "MSRV")


def run_c3(args):
    """The source-based objectives on the 2M-vertex workload: column nets with
    the diagonal plus 12 uniform off-diagonals (26M pins), vertex weights
    U{1..8}, K = 1024, eps = 0.05, 15 passes of each of TM, MSM, MSV, MSRV from
    its own initial stream, reference round schedules (TM one round per pass,
    the others rounds of --c3-round vertices). One line per run with a record
    per objective;
    `value` = pin-visits/s over the four objectives' passes together."""
    from fastpartitioner_amd import hypergraph as H
    from fastpartitioner_amd.engine import SCHED_REF, Engine, schedule

    K, eps, passes = 1024, 0.05, args.steps
    t = time.time()
    hg = H.workload("c3", seed=1, scale=args.scale)
    W = hg.total_weight()
    C, max_pw, min_pw = H.ref_constants(W, K, eps)
    parts = {}
    for mode in sorted(set(H.INIT_OF[m] for m in C3_OBJECTIVES)):
        parts[mode] = H.initial_partition_mode(hg, mode, K, C, max_pw, min_pw)
    print(f"# c3: nov={hg.nov} pins={hg.pins} K={K} prep {time.time() - t:.1f}s", file=sys.stderr, flush=True)
    per = {}
    total_s = 0.0
    with Engine(0) as eng:
        eng.upload(hg)
        for metric in C3_OBJECTIVES:
            part = parts[H.INIT_OF[metric]]
            sched = schedule(SCHED_REF, hg.nov if metric == "TM" else args.c3_round, C, max_pw, min_pw)
            if args.warmup:
                eng.set_partition(K, part)
                eng.refine(metric, args.warmup, sched)
            eng.set_partition(K, part)
            m0 = eng.metrics(vectors=True)
            eng.set_kernel_timing(True)
            t0 = time.perf_counter()
            st = eng.refine(metric, passes, sched)
            t1 = time.perf_counter()
            launches, eval_ms = eng.kernel_timing()
            eng.set_kernel_timing(False)
            m1 = eng.metrics()
            total_s += t1 - t0
            lam = lambda_visits(hg, m0["lam"])
            # evaluation bytes per launch (rounds of R vertices): the volume
            # evaluation reads the lists twice (12 B/vertex + 12 B/incidence +
            # 16 B/entry); the connection evaluation 12 B/vertex + 8 B/incidence
            # + 2 B/entry (its bitmap words are per gaining vertex, not counted)
            if metric in ("MSV", "MSRV"):
                eval_bytes = (12 * hg.nov + 12 * hg.pins + 16 * lam) * passes
            else:
                eval_bytes = (12 * hg.nov + 8 * hg.pins + 2 * lam) * passes
            per[metric] = {
                "ms_per_pass": round((t1 - t0) * 1e3 / passes, 3),
                "pin_visits_per_s": round(passes * hg.pins / (t1 - t0), 1),
                "moves": int(st.moves), "rounds": int(st.rounds),
                "initial": int(m0[metric]), "refined": int(m1[metric]),
                "eval_kernel": {"launches": int(launches), "avg_launch_us": round(eval_ms * 1e3 / max(launches, 1), 2),
                                "achieved_gbs": round(eval_bytes / max(eval_ms * 1e-3, 1e-9) / 1e9, 1),
                                "frac": round(eval_bytes / max(eval_ms * 1e-3, 1e-9) / 1e9 / HBM_PEAK_GBS, 4)},
            }
    cpu = None
    if not args.no_cpu:
        cpu = c3_cpu_baseline(args, K)
    value = len(C3_OBJECTIVES) * passes * hg.pins / total_s
    out = {
        "metric": "refine pin-visits/s (2M-vtx, K=1024, TM+MSM+MSV+MSRV)",
        "value": round(value, 1),
        "unit": "pin-visits/s",
        "n_gpus": 1,
        "steps": passes,
        "warmup": args.warmup,
        "ms_per_step": round(total_s * 1e3 / (passes * len(C3_OBJECTIVES)), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "colnet-uniform12-2M-K1024-w1to8", "nov": hg.nov, "pins": hg.pins, "K": K,
                   "eps": eps, "passes": passes, "round": args.c3_round, "objectives": per},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def c3_cpu_baseline(args, K):
    """The reference's own partition_par for TM, MSM, MGV (MSV) and MSRV, built
    from its sources (oracle/_ref), on a bounded sample: the same generator at
    200,000 vertices (the reference has no vertex weights: unit weights), one
    refinement pass, the box's CPU share of threads. Its initial streams at
    K = 1024 (the TM/MSM greedy costs O(K) per vertex and net) bound the
    sample size."""
    from fastpartitioner_amd import hypergraph as H
    from oracle import oracle as O

    info = cpu_info()
    threads = min(CPU_SHARE, info["affinity"], info["physical_cores"] or CPU_SHARE)
    hg = H.colnet(200_000, H.DEG_FIXED, 12, seed=1)
    res = {"sample": f"colnet diag+12, 200,000 vertices ({hg.pins} pins), unit weights, K={K}, 1 pass, "
                     f"{threads} OpenMP threads", "cores": threads, "cpu_model": info["model"], "kind": "reference",
           "unit": "pin-visits/s"}
    for metric, binary in (("TM", "tm"), ("MSM", "msm"), ("MSV", "mgv"), ("MSRV", "msrv")):
        if not O.ref_available(binary):
            res[metric] = {"error": f"oracle/_ref/ref_{binary} not built"}
            continue
        # partition_par races like TV's (its per-thread counts are read
        # before every thread wrote them) and can crash: a crashed run is
        # retried, then the thread count is halved
        crashes, t = [], threads
        while t >= 1 and metric not in res:
            for _ in range(2):
                env = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores", "OMP_NUM_THREADS": str(t)}
                try:
                    r = O.run_reference(binary, hg, K, 1, t, timeout=600, env_extra=env)
                except Exception as e:
                    crashes.append(f"{t} threads: " + ("segfault" if "rc=-11" in str(e) else str(e)[-60:]))
                    continue
                secs = r["refine_seconds"]
                res[metric] = {"refine_s": secs, "threads": t,
                               "pin_visits_per_s": round(hg.pins / secs, 1) if secs else None}
                if crashes:
                    res[metric]["crashed_runs_before"] = crashes
                break
            t //= 2
        if metric not in res:
            res[metric] = {"error": f"crashed at every thread count: {crashes}"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--round", type=int, default=262144, help="vertices per round")
    ap.add_argument("--sched", choices=("bal", "ref"), default="bal")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the 1M-vertex workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the reference CPU baseline")
    ap.add_argument("--cpu-passes", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", choices=("tv1m", "c3", "batch", "serve"), default="tv1m",
                    help="tv1m: the 1M-vertex TV refinement (default); c3: TM/MSM/MSV/MSRV on the 2M-vertex "
                         "K=1024 workload; batch: batched vs one-at-a-time refinement; serve: the serving loop "
                         "under Poisson arrivals")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM traffic measurement")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--instances", type=int, default=512, help="batch: number of hypergraphs")
    ap.add_argument("--passes", type=int, default=10, help="batch: passes per hypergraph")
    ap.add_argument("--tune-passes", type=int, default=3, help="batch: passes that double as autotuner trials")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--c3-round", type=int, default=262144,
                    help="c3: round size of the MSM, MSV and MSRV schedules (TM: one round per pass); "
                         "16384 gave the same refined objectives at half the rate")
    ap.add_argument("--lanes", type=int, default=16, help="batch: streams the members are spread over")
    ap.add_argument("--hw-queues", type=int, default=4, help="batch/serve: hardware queues of the process "
                    "(4, the HIP runtime's default, measured fastest for the batch)")
    ap.add_argument("--graphs", type=int, default=32, help="serve: resident hypergraphs")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    ap.add_argument("--workers", type=int, default=16, help="serve: worker threads")
    ap.add_argument("--replicas", type=int, default=1, help="serve: contexts per hypergraph at most")
    ap.add_argument("--budget-gib", type=int, default=64, help="serve: HBM budget of the requests in flight")
    ap.add_argument("--check", type=int, default=16, help="serve: requests rerun alone for the mix check")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES three times (HBM counter calibration)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        src = torch.ones(args.copy_calibration // 4, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration, "ok": bool((dst[-1] == 1).item())}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return
    if args.workload == "c3":
        if args.steps == 20 and "--steps" not in sys.argv:
            args.steps = 15
        run_c3(args)
        return
    run_tv1m(args)


if __name__ == "__main__":
    main()
