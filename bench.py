"""Benchmark of the one-GPU radix hash join (prints ONE JSON line).

Workload (default): R = S = 640M tuples (uint32 key + uint32 row-index value,
5.1 GB per relation), R keys distinct in [1, 2^31), 10% of S drawn from R, the
rest distinct in [2^31, 2^32); two-level radix with 14 total bits and
partition_size 512; blocked Bloom filter at 10 bits/key with k = 3 built from R
to filter S.  Inputs are generated on the device and are resident in HBM before
the timed region.  A step is one full join of the two relations.

value = (N_R + N_S) / (wall seconds per step), timed over K steps between
torch.cuda.synchronize() calls.  roofline = the dominant kernel (the kernel
function with the most device time per step, its launches summed, as rocprofv3
--stats ranks kernels): its algorithmic bytes per launch / its average launch
duration (HIP events on the library's stream, over the timed steps) against
8 TB/s; roofline.by_kernel gives the same for every kernel with algorithmic
bytes.  cpu_baseline = the reference's own CPU join
(src/sanity_check.cpp compiled into oracle/_ref), or the oracle's restatement
of it, on the benchmark's own inputs copied to the host, all host threads; the
oracle's row-emitting restatement is timed beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "brig-join_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (N_R, N_S, match ratio, zipf_s, params)
    "join640m_bf10": (640_000_000, 640_000_000, 0.10, 0.0,
                      dict(partition_size=512, radix_bits=14, use_bf=1, bf_bits_per_key=10.0,
                           num_hashes=3)),
    "join16m_nobf": (16_000_000, 16_000_000, 0.10, 0.0, dict(partition_size=1024)),
    "join640m_nobf": (640_000_000, 640_000_000, 0.10, 0.0, dict(partition_size=512, radix_bits=14)),
    "join64m_1280m_zipf": (64_000_000, 1_280_000_000, 0.01, 1.0,
                           dict(use_bf=1, bf_bits_per_key=12.0, num_hashes=4)),
    # batched form: 256 joins, |R| log-uniform in [64K, 16M], |S| uniform in [1, 4] x |R|,
    # match ratio uniform in [1%, 50%], join i generated with seed i
    "batch256": (None, None, None, 0.0, dict(joins=256)),
    # serving form: 8 resident R of 16M..320M keys (geometric), requests with a fresh S
    # of log-uniform size in [1M, 256M], 10% match, Poisson arrivals at 70% of the
    # single-join path's capacity
    "serve8": (None, None, 0.10, 0.0, dict(relations=8, pool=16, load=0.7)),
}


def algorithmic_bytes(kernel: str, n_r: int, n_s: int, m: int, m_s: int, bf_bytes: int,
                      bloom: bool, bf_stage: int = 1) -> float:
    """Compulsory streamed bytes of one launch of a kernel:
    a partition pass reads and writes 8 B per tuple, build+probe read 8 B per R
    and per S' tuple and write 12 B per row, the Bloom build reads 4 B per R key
    and writes the filter once, the Bloom filter reads 8 B per S tuple and writes
    8 B per matching S tuple; histogram passes read 4 B per key.  With the
    filter after the first-level pass (bf_stage 2), S's histogram and first pass
    cover all of S and the filter kernel also places the survivors (no second
    pass over S')."""
    s_part = m_s if bloom else n_s
    s_first = n_s if (bloom and bf_stage == 2) else s_part
    s_second = 0.0 if (bloom and bf_stage == 2) else 16.0 * s_part
    return {
        "hist_r": 4.0 * n_r, "scatter1_r": 16.0 * n_r, "scatter2_r": 16.0 * n_r,
        "bf_build": 4.0 * n_r + bf_bytes, "bf_filter": 8.0 * n_s + 8.0 * m_s,
        "hist_s": 4.0 * s_first, "scatter1_s": 16.0 * s_first, "scatter2_s": s_second,
        "probe": 8.0 * n_r + 8.0 * s_part + 12.0 * m,
    }.get(kernel, 0.0)


# Timing slot -> kernel function (rocprofv3's names without template arguments).
# The filter slot runs k_filter_l1 after S's first pass (stage 2), else
# k_bloom_prefilter.
KERNEL_OF_SLOT = {"hist_r": "k_hist", "hist_s": "k_hist", "scatter1_r": "k_scatter", "scatter2_r": "k_scatter",
                  "scatter1_s": "k_scatter", "scatter2_s": "k_scatter", "bf_build": "k_bloom_build",
                  "probe": "k_probe"}


def kernel_rooflines(kern_avg: dict, ab, bf_stage: int, workload: str) -> dict:
    """Per kernel function: its launches per step (slots with algorithmic bytes),
    device ms per step, algorithmic bytes and HBM traffic per launch, and the
    achieved rate (algorithmic bytes over device time) against 8 TB/s."""
    groups = {}
    for slot, ms in kern_avg.items():
        a = ab(slot)
        if a <= 0 or ms <= 0:
            continue
        fn = KERNEL_OF_SLOT.get(slot) or ("k_filter_l1" if bf_stage == 2 else "k_bloom_prefilter")
        g = groups.setdefault(fn, {"slots": [], "ms": 0.0, "alg": 0.0, "traffic": 0, "have_traffic": True})
        g["slots"].append(slot)
        g["ms"] += ms
        g["alg"] += a
        t = load_traffic(workload, slot)
        if t is None:
            g["have_traffic"] = False
        else:
            g["traffic"] += t
    out = {}
    for fn, g in groups.items():
        n = len(g["slots"])
        achieved = g["alg"] / (g["ms"] * 1e-3) / 1e9
        out[fn] = {"launches": n, "slots": g["slots"], "step_ms": round(g["ms"], 4),
                   "avg_ms": round(g["ms"] / n, 4), "alg_bytes": g["alg"] / n, "achieved": round(achieved, 1),
                   "frac": round(achieved / HBM_PEAK_GBS, 4),
                   "traffic": int(g["traffic"] / n) if g["have_traffic"] else None}
    return out


def total_algorithmic_bytes(n_r, n_s, m, m_s, bf_bytes, bloom) -> float:
    """Whole-join algorithmic bytes with two partition passes per relation: 40 B per
    tuple without the filter; with it, 44 B per R tuple plus the filter, 8 B per S
    key tested, 48 B per surviving S tuple, and 12 B per output row."""
    if bloom:
        return 44.0 * n_r + bf_bytes + 8.0 * n_s + 48.0 * m_s + 12.0 * m
    return 40.0 * (n_r + n_s) + 12.0 * m


def host_info() -> dict:
    """CPU model and NUMA node count of the host (Linux /proc and /sys)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nodes = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node")])
    except OSError:
        nodes = 0
    return {"cpu_model": model, "numa_nodes": nodes, "online_cpus": os.cpu_count()}


def cpu_baseline(rk, rv, sk, sv, m: int, budget_s: float = 60.0) -> dict:
    """The CPU join on the benchmark's own inputs (host copies of the device relations).

    value: the reference's own run_sanity_check (src/sanity_check.cpp, compiled
    from source into oracle/_ref), count only, all host threads: one warm-up, then
    the median of up to 3 runs within budget_s.  Without oracle/_ref the oracle's
    faithful restatement of it is timed instead (kind "port").  rows_port: the
    oracle's restatement extended to (key, r_val, s_val) rows, timed once."""
    import numpy as np
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    n_in = len(rk) + len(sk)
    if O.ref_available():
        kind = "reference"
        run = lambda: O.ref_sanity_count(rk, sk)  # noqa: E731
    else:
        kind = "port"

        def run():
            t0 = time.perf_counter()
            c = O.sanity_count(rk, sk, threads)
            return c, time.perf_counter() - t0
    t_start = time.perf_counter()
    c, warm = run()
    if c != m:
        raise RuntimeError(f"CPU join count {c} != {m}")
    times = []
    while len(times) < 3 and (time.perf_counter() - t_start) + warm < budget_s:
        c, sec = run()
        times.append(sec)
    if not times:
        times = [warm]
    med = float(np.median(times))
    t0 = time.perf_counter()
    n, _, _, _ = O.cpu_join(rk, rv, sk, sv, cap=m, threads=threads)
    rows_s = time.perf_counter() - t0
    if n != m:
        raise RuntimeError(f"CPU row join count {n} != {m}")
    mr, ms = len(rk) // 1_000_000, len(sk) // 1_000_000
    return {"value": n_in / med, "unit": "input tuples/s", "cores": threads, "kind": kind,
            "sample": f"the benchmark's own inputs ({mr}M x {ms}M tuples, copied from HBM), "
                      f"count-only join, median of {len(times)} after a warm-up; {med:.3f} s per join",
            **host_info(),
            "rows_port": {"value": n_in / rows_s, "seconds": round(rows_s, 3), "rows": int(n),
                          "what": "oracle restatement with (key, r_val, s_val) rows, one run"}}


def batch_sizes(n_joins: int, seed: int = 0):
    """Sizes and match ratios of the batched workload (deterministic)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_joins):
        nr = int(np.exp(rng.uniform(np.log(64 * 1024), np.log(16 * 1024 * 1024))))
        ns = int(nr * rng.uniform(1.0, 4.0))
        ratio = float(rng.uniform(0.01, 0.50))
        out.append((nr, ns, ratio))
    return out


def run_batch(args, n_joins: int) -> dict:
    """The batched form: all joins in one graph replay vs the single-join path one at a time."""
    import numpy as np
    import torch
    import brigjoin as bj

    sizes = batch_sizes(n_joins)
    tot_r = sum(s[0] for s in sizes)
    tot_s = sum(s[1] for s in sizes)
    dev = torch.device("cuda:0")
    R = torch.empty((2, tot_r), dtype=torch.int32, device=dev)
    S = torch.empty((2, tot_s), dtype=torch.int32, device=dev)
    ctx = bj.Context(0)
    specs, plain, expect = [], [], []
    orr = ors = 0
    for i, (nr, ns, ratio) in enumerate(sizes):
        rk, rv = R[0, orr:orr + nr], R[1, orr:orr + nr]
        sk, sv = S[0, ors:ors + ns], S[1, ors:ors + ns]
        m = int(min(ns, max(0, round(ratio * ns))))
        ctx.gen_build(i, nr, rk, rv)
        ctx.gen_probe(i, i ^ 0xDEADBEEF, nr, ns, m, sk, sv)
        specs.append(bj.join_spec(rk, rv, sk, sv, r_size=nr, s_size=ns, match_hint=ratio))
        plain.append((bj.default_params(r_size=nr, s_size=ns, inputs_on_device=1), (rk, rv, sk, sv)))
        expect.append(m)
        orr += nr
        ors += ns
    torch.cuda.synchronize()
    print(f"[bench] batch of {n_joins} joins generated; autotuning", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    tpath = getattr(args, "tune_table", None)
    if tpath and os.path.exists(tpath):
        # a table saved by an earlier run: profiled runs replay the configuration
        # of the timed one (the tuner's timings shift under counter collection)
        with open(tpath) as f:
            saved = json.load(f)
        entries = []
        for d in saved:
            e = bj.BjTuneEntry()
            for k, v in d.items():
                setattr(e, k, v)
            entries.append(e)
        bj.tune_set(ctx, entries)
        tuned = -1
    else:
        tuned = bj.autotune(ctx, specs)
    tune_s = time.perf_counter() - t0
    table = bj.tune_get(ctx)
    if tpath and tuned >= 0:
        with open(tpath, "w") as f:
            json.dump([{k: getattr(e, k) for k, _ in bj.BjTuneEntry._fields_} for e in table], f)
    print(f"[bench] autotuned {tuned} classes in {tune_s:.1f} s; graph replays", file=sys.stderr, flush=True)

    batch = bj.Batch(ctx, specs, use_graph=True)
    for _ in range(args.warmup):
        outs, _ = batch.run()
    torch.cuda.synchronize()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs, ms = batch.run()
        dev_ms.append(ms)
    wall = time.perf_counter() - t0
    bad = [i for i, o in enumerate(outs) if o.match_count != expect[i]]
    if bad:
        raise RuntimeError(f"batched join counts differ from the closed form at joins {bad[:8]}")
    ms_step = wall / args.steps * 1e3
    value = (tot_r + tot_s) / (wall / args.steps)

    print("[bench] single-join path", file=sys.stderr, flush=True)
    # the single-join path, one join at a time (same inputs, its own automatic
    # plan): timed without events between kernels, then once more with them for
    # its device time
    for p, _ in plain:
        p.collect_phase_times = 0
    # every join once first: the context's buffers grow to the largest join here,
    # not inside the timed loop (a grown buffer is a hipFree + hipMalloc)
    for p, d in plain:
        ctx.join(p, *d)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, (p, d) in enumerate(plain):
        r = ctx.join(p, *d)
        if r.match_count != expect[i]:
            raise RuntimeError(f"single-join count differs at join {i}")
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t0
    plain_value = (tot_r + tot_s) / plain_wall
    plain_dev = 0.0
    for p, d in plain:
        p.collect_phase_times = 1
        plain_dev += ctx.join(p, *d).phase_ms[6]

    # whole-batch algorithmic bytes (per join: the two-pass formula of its configuration)
    alg = 0.0
    for (nr, ns, ratio), o in zip(sizes, outs):
        m = int(o.match_count)
        f = int(round(nr * o.bf_bits_per_key / 8)) if o.bf_k else 0
        alg += total_algorithmic_bytes(nr, ns, m, m, f, bool(o.bf_k))
    dms = float(np.median(dev_ms))
    achieved = alg / (dms * 1e-3) / 1e9
    batch.close()
    ctx.close()
    return {
        "metric": "input tuples/s at 640M⋈640M on 1 MI355X; % HBM roofline; batched vs plain",
        "value": value, "unit": "input tuples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (device generator per join, seed = join index)",
        "config": {"workload": "batch256", "joins": n_joins, "tuples_r": tot_r, "tuples_s": tot_s,
                   "graph": True, "autotuned_classes": tuned if tuned >= 0 else f"table loaded from {tpath}",
                   "filter_stage_joins": {str(st): sum(1 for o in outs if o.bf_stage == st) for st in (0, 1, 2)}},
        "roofline": {"bound": "hbm", "kernel": "whole batch (graph replay)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic("batch256", "whole batch"), "alg_bytes": alg, "avg_ms": round(dms, 4)},
        "batched_vs_plain": {"plain_value": plain_value, "plain_ms": plain_wall * 1e3,
                             "plain_device_ms": plain_dev, "speedup": round(value / plain_value, 3)},
        "autotune": {"seconds": round(tune_s, 2),
                     "table": [{"kind": e.kind, "lg": e.lg, "sr": e.sr, "sel": e.sel,
                                "radix_bits": e.radix_bits, "tile": e.tile,
                                "bf_bits_per_key": e.bf_bits_per_key, "ms": round(e.ms, 4)}
                               for e in table]},
    }


def run_serve(args, opts: dict) -> dict:
    """The serving form under Poisson arrivals vs the single-join path per request."""
    import numpy as np
    import torch
    import brigjoin as bj

    rng = np.random.default_rng(1)
    nrel, npool, load = opts["relations"], opts["pool"], opts["load"]
    rel_sizes = [int(round(16_000_000 * (20.0 ** (i / (nrel - 1))))) for i in range(nrel)]
    dev = torch.device("cuda:0")
    ctx = bj.Context(0)
    R = []
    for i, n in enumerate(rel_sizes):
        rk = torch.empty(n, dtype=torch.int32, device=dev)
        rv = torch.empty(n, dtype=torch.int32, device=dev)
        ctx.gen_build(1000 + i, n, rk, rv)
        R.append((rk, rv, n))
    pool = []
    for j in range(npool):
        ri = int(rng.integers(0, nrel))
        ns = int(np.exp(rng.uniform(np.log(1 << 20), np.log(256 << 20))))
        sk = torch.empty(ns, dtype=torch.int32, device=dev)
        sv = torch.empty(ns, dtype=torch.int32, device=dev)
        m = int(round(0.10 * ns))
        ctx.gen_probe(1000 + ri, 5000 + j, rel_sizes[ri], ns, m, sk, sv)
        pool.append((ri, sk, sv, ns, m))
    torch.cuda.synchronize()

    # the single-join path per request: partitions R and S, filter from R, probe
    plain_ms = []
    for ri, sk, sv, ns, m in pool:
        rk, rv, nr = R[ri]
        p = bj.default_params(r_size=nr, s_size=ns, inputs_on_device=1, use_bf=1, bf_bits_per_key=10.0,
                              num_hashes=3, collect_phase_times=0)
        ctx.join(p, rk, rv, sk, sv)  # warm
        t0 = time.perf_counter()
        res = ctx.join(p, rk, rv, sk, sv)
        plain_ms.append((time.perf_counter() - t0) * 1e3)
        if res.match_count != m:
            raise RuntimeError("single-join count differs from the closed form")
    ctx.close()
    mean_plain = float(np.mean(plain_ms))
    lam = load / (mean_plain * 1e-3)

    srv = bj.Server(0, slots=opts.get("slots", 3))
    t0 = time.perf_counter()
    for rk, rv, n in R:
        srv.add_relation(rk, rv, n)
    torch.cuda.synchronize()
    reg_s = time.perf_counter() - t0
    alone_ms = []
    for ri, sk, sv, ns, m in pool:
        for _ in range(2):
            rid = srv.submit(ri, sk, sv, ns)
            r = srv.wait(rid)
            srv.release(rid)
        if r.status != 0 or r.match_count != m:
            raise RuntimeError("served count differs from the closed form")
        alone_ms.append((r.t_end - r.t_submit) * 1e3)

    # Poisson arrivals over the run; completed requests are polled and released
    duration = args.serve_seconds
    arrivals = []
    t = 0.0
    while True:
        t += float(rng.exponential(1.0 / lam))
        if t > duration:
            break
        arrivals.append((t, int(rng.integers(0, npool))))
    lat, tuples, s_tuples, done = [], 0, 0, 0
    bad = 0
    submit_at = {}
    start = time.perf_counter()
    k = 0
    last_progress = start
    while k < len(arrivals) or done < len(arrivals):
        now = time.perf_counter() - start
        while k < len(arrivals) and arrivals[k][0] <= now:
            ri, sk, sv, ns, m = pool[arrivals[k][1]]
            rid = srv.submit(ri, sk, sv, ns)
            submit_at[rid] = (arrivals[k][1], m)
            k += 1
        for r in srv.poll(512):
            j, m = submit_at.pop(r.id)
            if r.status != 0 or r.match_count != m:
                bad += 1
            lat.append((r.t_end - r.t_submit) * 1e3)
            tuples += pool[j][3] + rel_sizes[pool[j][0]]
            s_tuples += pool[j][3]
            done += 1
            srv.release(r.id)
        if time.perf_counter() - last_progress > 10.0:
            print(f"[serve] t={now:.1f}s submitted={k} done={done}", file=sys.stderr, flush=True)
            last_progress = time.perf_counter()
        nxt = arrivals[k][0] - (time.perf_counter() - start) if k < len(arrivals) else 0.0005
        time.sleep(max(0.0, min(nxt, 0.0005)))
    wall = time.perf_counter() - start

    # capacity: the same request mix submitted at once (about 10 s of the single-join
    # path's work), the server admitting and overlapping as slots and arena allow
    sat_n = max(len(pool), int(10.0 / (mean_plain * 1e-3)))
    sat_jobs = [arrivals[i % len(arrivals)][1] if arrivals else i % len(pool) for i in range(sat_n)]
    print(f"[serve] saturation: {sat_n} requests at once", file=sys.stderr, flush=True)
    sat_ids = {}
    t_sat = time.perf_counter()
    for j in sat_jobs:
        ri, sk, sv, ns, m = pool[j]
        sat_ids[srv.submit(ri, sk, sv, ns)] = (j, m)
    sat_tuples, sat_done = 0, 0
    while sat_done < sat_n:
        for r in srv.poll(512):
            j, m = sat_ids.pop(r.id)
            if r.status != 0 or r.match_count != m:
                bad += 1
            sat_tuples += pool[j][3] + rel_sizes[pool[j][0]]
            sat_done += 1
            srv.release(r.id)
        time.sleep(0.0002)
    sat_wall = time.perf_counter() - t_sat
    sat_plain_s = sum(plain_ms[j] for j in sat_jobs) * 1e-3
    st = srv.stats()
    srv.close()
    if bad:
        raise RuntimeError(f"{bad} served requests had a wrong count")

    # the single-join path under the same arrivals: one FIFO server with the measured
    # per-request times (Lindley recursion)
    free_at, plain_lat = 0.0, []
    for t_arr, j in arrivals:
        begin = max(t_arr, free_at)
        free_at = begin + plain_ms[j] * 1e-3
        plain_lat.append((free_at - t_arr) * 1e3)
    tot_plain_tuples = sum(pool[j][3] + rel_sizes[pool[j][0]] for _, j in arrivals)
    lat = np.array(lat)
    value = tuples / wall
    return {
        "metric": "input tuples/s at 640M⋈640M on 1 MI355X; % HBM roofline; batched vs plain",
        "value": value, "unit": "input tuples/s", "n_gpus": 1, "steps": len(arrivals),
        "warmup": len(pool), "ms_per_step": float(np.mean(lat)), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (device generator; S pool reused across requests)",
        "config": {"workload": "serve8", "relations": rel_sizes, "pool_s_sizes": [p[3] for p in pool],
                   "arrival_rate_per_s": round(lam, 2), "offered_load_vs_plain": load,
                   "seconds": duration, "slots": opts.get("slots", 3)},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "mean": round(float(lat.mean()), 3), "max": round(float(lat.max()), 3)},
        "served": {"requests": int(len(lat)),
                   "request_input_tuples_per_s": value,
                   "request_input_tuples": "per request |S| + |R| of its resident relation, the join's input "
                                           "as the single-join path counts it (R is partitioned once, at "
                                           "registration, not per request)",
                   "s_tuples_per_s": s_tuples / wall, "wall_s": round(wall, 2),
                   "max_in_flight": int(st.max_in_flight), "arena_peak_bytes": int(st.arena_peak_bytes),
                   "relation_bytes": int(st.relation_bytes), "register_s": round(reg_s, 3),
                   "alone_ms_mean": round(float(np.mean(alone_ms)), 3)},
        "saturated": {"requests": sat_n, "wall_s": round(sat_wall, 3),
                      "request_input_tuples_per_s": sat_tuples / sat_wall,
                      "vs_plain_same_requests": round(sat_plain_s / sat_wall, 3)},
        "plain_per_request": {"service_ms_mean": round(mean_plain, 3),
                              "capacity_tuples_per_s": tot_plain_tuples / (sum(plain_ms[j] for _, j in arrivals) * 1e-3),
                              "p50_ms_same_arrivals": round(float(np.percentile(plain_lat, 50)), 3),
                              "p99_ms_same_arrivals": round(float(np.percentile(plain_lat, 99)), 3)},
    }


def load_traffic(workload: str, kernel: str):
    """HBM bytes per launch of a kernel slot, from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes recorded in profiles/traffic_<workload>.json (each counter scaled by what it
    reads for a plain device copy of known size); None when not recorded."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    e = d.get(kernel)
    return int(e["bytes"]) if isinstance(e, dict) and "bytes" in e else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="join640m_bf10", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-bits-sweep", action="store_true",
                    help="skip the joins at one radix bit fewer and more (profiled runs: one configuration)")
    ap.add_argument("--cpu-budget-s", type=float, default=60.0,
                    help="wall-clock budget of the timed CPU count-only runs")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="serve8: length of the arrival process")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--probe-breakdown", action="store_true",
                    help="one more join with the probe's in-kernel phase timers (probe_phases_ms)")
    ap.add_argument("--bf-stage", type=int, default=None, help="override: 1 before, 2 after pass 1")
    ap.add_argument("--pass1-bits", type=int, default=None, help="override the first scatter pass bits")
    ap.add_argument("--s-pass1-bits", type=int, default=None, help="override S's first scatter pass bits")
    ap.add_argument("--scatter-wg", type=int, default=None, help="scatter workgroup size (512 or 1024)")
    ap.add_argument("--tile", type=int, default=None, help="probe LDS table capacity (4096 or 8192)")
    ap.add_argument("--host-inputs", action="store_true",
                    help="R and S in pinned host memory: the join copies them (value is then the "
                         "PCIe-inclusive rate, not the HBM-resident one)")
    ap.add_argument("--s-chunks", type=int, default=None,
                    help="join S in this many chunks (host chunks copied while the previous one is joined)")
    ap.add_argument("--chunk-probe", type=int, default=None,
                    help="with --s-chunks: 0 one probe after the last chunk (pooled S'), 1 one per chunk")
    ap.add_argument("--tune-table", default=None,
                    help="batch256: load the autotuner table from this JSON file if it exists, else "
                         "autotune and save it there")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="only run a plain device copy of this many GiB (calibrates rocprof "
                         "FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures the one-GPU join only: run it with --gpus 1 (no multi-GPU path "
              "is built)", file=sys.stderr)
        return 2

    import torch
    import brigjoin as bj

    if args.copy_probe_gib > 0:
        nbytes = int(args.copy_probe_gib * 2**30) // 4 * 4
        src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nbytes, "launches": 3}), flush=True)
        return 0

    n_r, n_s, ratio, zipf_s, params = WORKLOADS[args.workload]
    if args.workload == "batch256":
        print(json.dumps(run_batch(args, params["joins"])), flush=True)
        return 0
    if args.workload == "serve8":
        print(json.dumps(run_serve(args, params)), flush=True)
        return 0
    m_s = int(min(n_s, max(0, round(ratio * n_s))))
    dev = torch.device("cuda:0")
    rk = torch.empty(n_r, dtype=torch.int32, device=dev)
    rv = torch.empty(n_r, dtype=torch.int32, device=dev)
    sk = torch.empty(n_s, dtype=torch.int32, device=dev)
    sv = torch.empty(n_s, dtype=torch.int32, device=dev)
    ctx = bj.Context(0)
    ctx.gen_build(42, n_r, rk, rv)
    ctx.gen_probe(42, 42 ^ 0xDEADBEEF, n_r, n_s, m_s, sk, sv, zipf_s=zipf_s)
    torch.cuda.synchronize()
    if args.bf_stage is not None:
        params = dict(params, bf_stage=args.bf_stage)
    if args.pass1_bits is not None:
        params = dict(params, pass1_bits=args.pass1_bits)
    if args.s_pass1_bits is not None:
        params = dict(params, s_pass1_bits=args.s_pass1_bits)
    if args.scatter_wg is not None:
        params = dict(params, scatter_wg=args.scatter_wg)
    if args.tile is not None:
        params = dict(params, tile=args.tile)
    if args.s_chunks is not None:
        params = dict(params, s_chunks=args.s_chunks)
    if args.chunk_probe is not None:
        params = dict(params, chunk_probe=args.chunk_probe)
    if args.host_inputs:
        # pinned host copies of the generated relations (the device ones are freed)
        def pinned(t):
            h = torch.empty(t.numel(), dtype=t.dtype, pin_memory=True)
            h.copy_(t)
            return h
        rk, rv, sk, sv = (pinned(t) for t in (rk, rv, sk, sv))
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    p = bj.default_params(r_size=n_r, s_size=n_s, inputs_on_device=0 if args.host_inputs else 1, **params)
    bloom = bool(params.get("use_bf"))

    # The timed steps record no events between kernels (each costs the device
    # a few microseconds); the per-kernel times and the roofline come from the
    # same number of instrumented joins after them.
    p.collect_phase_times = 0
    for _ in range(args.warmup):
        res = ctx.join(p, rk, rv, sk, sv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = ctx.join(p, rk, rv, sk, sv)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if res.match_count != m_s:
        print(f"wrong match count {res.match_count} != {m_s}", file=sys.stderr)
        return 1
    p.collect_phase_times = 1
    kern_sum = {k: 0.0 for k in bj.KERNELS}
    total_dev = 0.0
    h2d_sum = 0.0
    for _ in range(args.steps):
        res = ctx.join(p, rk, rv, sk, sv)
        for k, v in res.kernels().items():
            kern_sum[k] += v
        total_dev += res.phase_ms[6]
        h2d_sum += res.phase_ms[0]
    torch.cuda.synchronize()
    if res.match_count != m_s:
        print(f"wrong match count {res.match_count} != {m_s}", file=sys.stderr)
        return 1

    # The same join at one radix bit fewer and more than configured (the
    # configured run above is the headline; this is the autotuned-bits report)
    bits_sweep = None
    if (not args.no_bits_sweep and not args.host_inputs and not args.s_chunks and args.pass1_bits is None
            and args.s_pass1_bits is None):
        base_bits = int(res.bits1) + int(res.bits2)
        cand = {}
        for b in (base_bits - 1, base_bits, base_bits + 1):
            if b < 1 or b > 15:
                continue
            q = bj.default_params(r_size=n_r, s_size=n_s, inputs_on_device=1, **dict(params, radix_bits=b))
            q.collect_phase_times = 0
            try:
                r = ctx.join(q, rk, rv, sk, sv)
            except bj.BjError:
                continue
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            reps = max(1, min(3, args.steps))
            for _ in range(reps):
                r = ctx.join(q, rk, rv, sk, sv)
            torch.cuda.synchronize()
            if r.match_count != m_s:
                print(f"wrong match count at {b} radix bits", file=sys.stderr)
                return 1
            cand[b] = (time.perf_counter() - t1) / reps * 1e3
        if cand:
            best = min(cand, key=cand.get)
            bits_sweep = {"configured_bits": base_bits, "ms_by_bits": {str(b): round(v, 3) for b, v in cand.items()},
                          "best_bits": best, "best_ms": round(cand[best], 3),
                          "best_value": (n_r + n_s) / (cand[best] * 1e-3)}

    probe_phases = None
    if args.probe_breakdown:
        p.probe_breakdown = 1
        res = ctx.join(p, rk, rv, sk, sv)
        probe_phases = {k: round(v, 4) for k, v in res.probe_phases().items()}
        probe_phases["probe_ms"] = round(res.kernels()["probe"], 4)
        p.probe_breakdown = 0

    ms_step = wall / args.steps * 1e3
    value = (n_r + n_s) / (wall / args.steps)
    kern_avg = {k: v / args.steps for k, v in kern_sum.items()}
    bf_bytes = int(res.bf_words) * 4
    m = int(res.match_count)
    ms_sv = int(res.s_survivors)
    stage = int(res.bf_stage)
    ab = lambda k: algorithmic_bytes(k, n_r, n_s, m, m_s, bf_bytes, bloom, stage)  # noqa: E731
    by_kernel = kernel_rooflines(kern_avg, ab, stage, args.workload)
    dominant = max(by_kernel, key=lambda k: by_kernel[k]["step_ms"])
    dk = by_kernel[dominant]
    per_kernel = {k: {"ms": round(v, 4),
                      "alg_GBps": round(ab(k) / (v * 1e-3) / 1e9, 1) if v > 0 else None}
                  for k, v in kern_avg.items() if v > 0}
    B = total_algorithmic_bytes(n_r, n_s, m, m_s, bf_bytes, bloom)

    out = {
        "metric": "input tuples/s at 640M⋈640M on 1 MI355X; % HBM roofline; batched vs plain",
        "value": value,
        "unit": "input tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (device generator: distinct R keys, exact match ratio)",
        "config": {"workload": args.workload, "n_r": n_r, "n_s": n_s, "match_ratio": ratio,
                   "inputs": "host, pinned (value includes the PCIe copies)" if args.host_inputs else "HBM",
                   "s_chunks": int(res.s_chunks),
                   "zipf_s": zipf_s, **{k: v for k, v in params.items()},
                   "bits1": int(res.bits1), "bits2": int(res.bits2), "tile": int(res.tile),
                   "bf_stage": int(res.bf_stage), "pass1_bits": int(res.pass1_bits),
                   "s_pass1_bits": int(res.s_pass1_bits),
                   "bf_seg_bits": int(res.bf_seg_bits), "bf_k": int(res.bf_k)},
        "roofline": {"bound": "hbm", "kernel": dominant, "launches": dk["launches"], "achieved": dk["achieved"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dk["frac"], "traffic": dk["traffic"],
                     "alg_bytes": dk["alg_bytes"], "avg_ms": dk["avg_ms"],
                     "step_share": round(dk["step_ms"] / (total_dev / args.steps), 4),
                     "by_kernel": by_kernel},
        "whole_join": {"alg_bytes": B, "device_ms": round(total_dev / args.steps, 4),
                       "alg_frac_of_8TBps": round(B / (total_dev / args.steps * 1e-3) / 8e12, 4),
                       "matches": m, "s_survivors": ms_sv,
                       "vlsplit_items": int(res.heavy_items),
                       "h2d_ms_before_join": round(h2d_sum / args.steps, 4)},
        "kernels": per_kernel,
    }
    if bits_sweep is not None:
        out["autotuned_bits"] = bits_sweep
    if probe_phases is not None:
        out["probe_phases_ms"] = probe_phases
    ctx.close()
    host = None
    if not args.no_cpu_baseline:
        import numpy as np
        host = [np.array(t.cpu().numpy().view(np.uint32)) for t in (rk, rv, sk, sv)]
    # Pinned host buffers go back to the HIP runtime now, while it is alive: left
    # to the end of the process, torch's pinned-memory cache frees them from a
    # static destructor, after the runtime (or a profiler's tool library) has
    # already been torn down.
    del rk, rv, sk, sv
    torch.cuda.synchronize()
    if args.host_inputs:
        torch._C._host_emptyCache()
    if host is not None:
        print(f"[bench] CPU baseline on {n_r} x {n_s} tuples", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(*host, m_s, args.cpu_budget_s)
        out["cpu_baseline"]["gpu_speedup"] = round(value / out["cpu_baseline"]["value"], 1)
        del host
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
