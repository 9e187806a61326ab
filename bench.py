"""Benchmark: SCC decomposition throughput on one MI355X.

A step is one full SCC decomposition of a device-resident graph: CSR and
reverse CSR already in HBM when the timed region starts, canonical labels in
HBM when it ends.  Everything the decomposition derives from the CSR (the
first-neighbour and degree-product arrays) is built inside the step.
Default workload: R-MAT scale 24 (16,777,216 vertices, 268,435,456 arcs,
a=0.57 b=0.19 c=0.19, seeded permutation), -a g with direction-optimised
expansion.

Printed as ONE JSON line:
  value        SCC edges/s = E * steps / wall time of the timed steps
  roofline     the dominant kernel, expand_do_kernel, over the first-pivot
               FW+BW search: `achieved` = the bytes of the work that search
               did (vertices expanded, arcs scanned, candidates and parents
               probed, claims: counted on the device in an untimed run with
               sizes, at fixed bytes per unit) over its device time (HIP
               events on the library's stream) in the timed steps, against
               8 TB/s HBM.  `model_bytes` is the top-down byte model
               12|R| + 8|E_R| per direction over the same time, for
               comparison; `traffic` is the search's HBM bytes measured with
               rocprof counters on this same workload (profiles/)
  phase_ms     per-phase device times (separate untimed runs that record a
               timestamp at every phase boundary), with E / t per phase
  cpu_baseline the oracle's multithreaded FW-BW-trim (vHong phase order) on
               the same graph, all host threads used (up to 16), and a
               single-thread Tarjan, with the host's core count and model
  extra keys   rmat22 (with the 64-run algorithm x warp x trim matrix),
               grid4096, er16k, batched (4,096-graph mix vs one graph per
               launch), serving (capacity, then 70% of it), partitioned
               (8,388,608 vertices, 200 planted communities, -a d)
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
WORKLOADS = {
    "rmat22": "R-MAT scale 22, 4,194,304 V, 67,108,864 arcs, a=.57 b=.19 c=.19, seed 22",
    "rmat24": "R-MAT scale 24, 16,777,216 V, 268,435,456 arcs, a=.57 b=.19 c=.19, seed 24",
    "er16k": "ER 16,384 V, 131,072 draws + 64 planted cycles (2..256), seed 1",
    "grid4096": "grid 4096x4096, arc probability 0.7, seed 4096",
    # measured alone (one JSON line each, value = graphs/s)
    "batch4096": "the 4096-graph mix, batched persistent form vs one graph per launch",
    "serve": "Poisson request stream into the continuous-batching server",
    "part200": "-a d on 8,388,608 V with 200 planted communities, partitioned by community",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def generate_on_device(ctx, name):
    """The workload generated in place on the device (bit-identical to the
    host generators of the same arguments); returns the graph's (V, E)."""
    if name == "rmat22":
        ctx.gen_rmat_dev(22, seed=22)
    elif name == "rmat24":
        ctx.gen_rmat_dev(24, seed=24)
    elif name == "er16k":
        ctx.gen_er_dev(16384, 131072, cycles=64, cmin=2, cmax=256, seed=1)
    elif name == "grid4096":
        ctx.gen_grid_dev(4096, 0.7, seed=4096)
    else:
        raise SystemExit(f"unknown workload {name}")
    return ctx.graph.V, ctx.graph.E


def model_bytes(nv, ne):
    """Top-down reachability byte model per direction: 12 B per reached
    vertex, 8 B per arc scanned out of the reached set."""
    return 12 * nv + 8 * ne


def profiled_traffic(workload):
    """HBM bytes of one first-pivot search of `workload`, measured with
    rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy), when a
    profile of that exact workload is committed."""
    p = os.path.join(ROOT, "profiles", f"reach_traffic_{workload}.json")
    if not os.path.exists(p):
        return None, None
    with open(p) as f:
        d = json.load(f)
    if d.get("workload") != WORKLOADS.get(workload):   # a profile of another configuration
        return None, None
    return d.get("hbm_bytes_per_search"), (f"profiles/reach_traffic_{workload}.md: rocprofv3 FETCH_SIZE + "
                                           "WRITE_SIZE over this workload's first search, copy-calibrated")


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(g, workload: str, threads: int, budget_s: float = 12.0):
    """The oracle's FW-BW-trim in the vHong phase order on `threads` host
    threads (median of the runs that fit `budget_s`, at least 3), plus one
    single-thread Tarjan run; both checked against each other."""
    from oracle import pyoracle as orc
    nproc, model = host_info()
    runs = []
    lab = None
    t_end = time.perf_counter() + budget_s
    while len(runs) < 3 or (time.perf_counter() < t_end and len(runs) < 15):
        t = time.perf_counter()
        lab, n, _ = orc.vhong(g, True, True, threads)
        runs.append(time.perf_counter() - t)
    runs.sort()
    med = runs[len(runs) // 2]
    t = time.perf_counter()
    tl, tn = orc.tarjan(g)
    t_tarjan = time.perf_counter() - t
    return {"value": g.E / med, "unit": "edges/s", "cores": threads, "kind": "port",
            "sample": f"{WORKLOADS[workload]} (the headline workload itself), oracle vHong phase order "
                      f"(trim-1, FW-BW, trim-1/2, min-label, pivot rounds), median of {len(runs)} runs",
            "ms_per_decomposition": 1e3 * med, "host_nproc": nproc, "host_cpu_model": model,
            "tarjan_single_thread_ms": 1e3 * t_tarjan, "tarjan_edges_per_s": g.E / t_tarjan,
            "labels_equal_tarjan": bool(np.array_equal(lab, tl)) and n == tn}


def bench_single(ctx, name, steps, warmup, algo="g", warp=1, expand=4, phases=True):
    """One workload through the single-graph engine: untimed counted run,
    warm-up, `steps` timed decompositions, per-phase breakdown."""
    import detectingscc_amd as scc
    t0 = time.perf_counter()
    V, E = generate_on_device(ctx, name)
    gen_s = time.perf_counter() - t0
    log(f"generated {name} on the device: V={V} E={E} in {gen_s:.2f}s")
    kw = dict(algo=algo, warp=warp, expand=expand, sizes=False)
    # untimed: |R|, |E_R|, the first search's work units, largest SCC, and
    # the labels every timed run must reproduce
    counted = ctx.run(**dict(kw, sizes=True))
    want = ctx.labels()
    hist, hist_ms = ctx.histogram()
    for i in range(warmup):
        st = ctx.run(**kw)
        assert st.num_sccs == counted.num_sccs, (st.num_sccs, counted.num_sccs)
        log(f"{name} warmup {i}: {st.ms_scc:.3f} ms device, {st.num_sccs} SCCs")
    stats = []
    t = time.perf_counter()
    for _ in range(steps):
        stats.append(ctx.run(**kw))     # synchronous: returns after the last kernel
    wall = time.perf_counter() - t
    assert all(s.num_sccs == counted.num_sccs for s in stats)
    assert np.array_equal(ctx.labels(), want), "timed run's labels differ from the counted run's"
    fr = sorted(s.ms_first_reach for s in stats)
    fr_ms = fr[len(fr) // 2]
    work = counted.first_work_bytes
    model = (model_bytes(counted.first_fw_vertices, counted.first_fw_arcs) +
             model_bytes(counted.first_bw_vertices, counted.first_bw_arcs))
    out = {
        "workload": WORKLOADS[name], "V": V, "E": E, "algo": algo, "warp": warp, "expand": expand,
        "generated_on_device_s": gen_s,
        "value": E * steps / wall, "ms_per_step": 1e3 * wall / steps,
        "device_ms_per_step": sum(s.ms_scc for s in stats) / len(stats),
        "num_sccs": counted.num_sccs, "largest_scc": counted.largest,
        "labels_checked": "every timed run's SCC count and the last run's labels equal the counted run's",
        "search_ms": fr_ms, "search_work_bytes": work,
        "search_GBps": work / (fr_ms * 1e6) if fr_ms > 0 else None,
        "search_work_units": dict(zip(["td_vertices", "td_arcs", "claims", "bu_swept", "bu_candidates",
                                       "bu_first_parents", "bu_rows", "bu_more_parents"],
                                      counted.first_work)),
        "search_model_bytes": model, "fw": [counted.first_fw_vertices, counted.first_fw_arcs],
        "bw": [counted.first_bw_vertices, counted.first_bw_arcs],
        "histogram_device_ms": hist_ms, "distinct_scc_sizes": len(hist),
    }
    traffic, src = profiled_traffic(name)
    if traffic:
        out["search_traffic_bytes"] = traffic
        out["search_traffic_GBps"] = traffic / (fr_ms * 1e6) if fr_ms > 0 else None
        out["search_traffic_source"] = src
    if phases:
        phased = [ctx.run(**dict(kw, phase_times=True)) for _ in range(max(1, min(steps, 5)))]
        ph = {k: sum(s.ms_phase[k] for s in phased) / len(phased) for k in scc.PHASES}
        out["phase_ms"] = {k: round(v, 4) for k, v in ph.items() if v > 0}
        out["phase_edges_per_s"] = {k: E / (v * 1e-3) for k, v in ph.items() if v > 0}
    return out


def bench_matrix(ctx, g):
    """The 64-run matrix on the uploaded R-MAT scale 22 graph: {g, h} x
    trims, {x, y} x W in {1..64} x trims, each with the reference's own
    mapping (vertex per lane; virtual warp of W lanes for the first search),
    then the 16 (algorithm, trims) runs with the automatic expansion a
    drop-in caller gets (direction-optimised; W does not apply).  Device ms
    of one decomposition each (after one warm-up run)."""
    import detectingscc_amd as scc
    n0 = None

    def one(a, w, p, q, expand):
        nonlocal n0
        ctx.run(algo=a, warp=w, trim1=p, trim2=q, expand=expand, sizes=False)
        st = ctx.run(algo=a, warp=w, trim1=p, trim2=q, expand=expand, sizes=False)
        n0 = st.num_sccs if n0 is None else n0
        assert st.num_sccs == n0, (a, w, p, q, expand, st.num_sccs, n0)
        return round(st.ms_scc, 3)

    runs = ([(a, 1, p, q) for a in "gh" for p in (1, 0) for q in (1, 0)] +
            [(a, w, p, q) for a in "xy" for w in (1, 2, 4, 8, 16, 32, 64) for p in (1, 0) for q in (1, 0)])
    res = {f"-a {a} -w {w} -p {p} -q {q}": one(a, w, p, q, scc.EXPAND_VWARP if a in "xy" else scc.EXPAND_VERTEX)
           for a, w, p, q in runs}
    dflt = {f"-a {a} -p {p} -q {q}": one(a, 1, p, q, scc.EXPAND_AUTO)
            for a in "ghxy" for p in (1, 0) for q in (1, 0)}
    return {"ms": res, "edges_per_s_best": g.E / (min(res.values()) * 1e-3),
            "edges_per_s_worst": g.E / (max(res.values()) * 1e-3),
            "default_expansion_ms": dflt,
            "default_expansion_edges_per_s_best": g.E / (min(dflt.values()) * 1e-3),
            "default_expansion_edges_per_s_worst": g.E / (max(dflt.values()) * 1e-3)}


def bench_batched(graphs=4096, steps=3, sample=1024, serial_baseline=True):
    import detectingscc_amd as scc
    with scc.Context(0) as ctx:
        t0 = time.perf_counter()
        ctx.gen_batch_mix_dev(graphs, 8, 16, 2.0, 16.0, seed=4096)
        gen_s = time.perf_counter() - t0
        b = ctx.batch_info()
        log(f"generated the {graphs}-graph mix on the device: {int(b.V.sum())} V, {int(b.E.sum())} arcs "
            f"in {gen_s:.2f}s")
        report = ctx.batch_autotune(sample=sample)
        ctx.batch_run(autotuned=True)                        # warm-up
        runs = [ctx.batch_run(autotuned=True) for _ in range(steps)]
        best = min(runs, key=lambda r: r.ms)
        labs, nscc = ctx.batch_labels()
        serial = None
        if serial_baseline:
            serial = ctx.batch_run_serial(expand=4)          # one graph per launch, DO expansion
            slabs, snscc = ctx.batch_labels()
    same = None
    if serial is not None:
        same = bool(np.array_equal(nscc, snscc)) and all(np.array_equal(x, y) for x, y in zip(labs, slabs))
        log(f"batched {best.ms:.2f} ms ({best.graphs_per_s:.0f} graphs/s), one-graph-per-launch "
            f"{serial.ms:.0f} ms ({serial.graphs_per_s:.0f} graphs/s), labels identical: {same}")
    return {
        "workload": f"{graphs} graphs, V log-uniform [2^8, 2^16], degree [2, 16], 40% ER / 40% R-MAT / "
                    "20% planted cycles 2..512, seed 4096",
        "vertices": int(b.V.sum()), "arcs": int(b.E.sum()), "generated_on_device_s": gen_s,
        "graphs_per_s": best.graphs_per_s, "ms": best.ms, "p50_ms": best.p50_ms, "p99_ms": best.p99_ms,
        "launches": best.launches, "steps": steps,
        "serial_graphs_per_s": serial.graphs_per_s if serial else None,
        "serial_ms": serial.ms if serial else None,
        "serial_p99_ms": serial.p99_ms if serial else None,
        "speedup_vs_one_graph_per_launch": best.graphs_per_s / serial.graphs_per_s if serial else None,
        "labels_identical": same,
        "autotune": report.strip().splitlines(),
    }


def _serve(c, graphs, gaps):
    """Submit `graphs` with inter-arrival `gaps` (None: all at once); wall
    time from the first submit to the last completion and per-request
    latencies (submit -> labels on the host)."""
    t = time.perf_counter()
    due = t
    ids = []
    for k, g in enumerate(graphs):
        if gaps is not None:
            due += gaps[k]
            while time.perf_counter() < due:
                pass
        ids.append(c.server_submit(g))
    c.server_wait_all()
    wall = time.perf_counter() - t
    lat = np.array([c.server_poll(rid)[2] for rid in ids])
    return wall, lat


def bench_serving(requests=400, budget_gib=32, slots=3, load=0.7):
    """Capacity first (every request submitted at once: completed / wall),
    then Poisson arrivals at `load` x capacity; p50/p99 at that load."""
    import detectingscc_amd as scc
    t0 = time.perf_counter()
    b = scc.gen_serving_mix(requests, 16384, 1.5, 1 << 22, 8.0, seed=8)
    graphs = [b.graph(i) for i in range(requests)]
    log(f"built {requests} requests: {int(b.V.sum())} V, {int(b.E.sum())} arcs, largest {int(b.V.max())} V "
        f"in {time.perf_counter() - t0:.1f}s")
    with scc.Context(0) as c:
        c.server_start(budget_bytes=budget_gib << 30, max_batch=256, slots=slots, solo_vertices=1 << 16,
                       lanes=16, block=1024)
        # warm-up (untimed): the largest requests, a few times, so every
        # slot's grow-only device buffers reach their size before the runs
        big = sorted(range(requests), key=lambda i: -graphs[i].E)[:slots]
        for _ in range(slots):
            _serve(c, [graphs[i] for i in big], None)
        cap_wall, _ = _serve(c, graphs, None)
        capacity = requests / cap_wall
        rate = load * capacity
        gaps = np.random.default_rng(1).exponential(1.0 / rate, size=requests)
        st0 = c.server_stats()
        wall, lat = _serve(c, graphs, gaps)
        st = c.server_stats()
        c.server_stop()
    p50, p99 = (float(x) for x in np.percentile(lat, [50, 99]))
    log(f"capacity {capacity:.0f} requests/s; at {rate:.0f}/s: p50 {p50:.2f} ms, p99 {p99:.2f} ms")
    return {
        "workload": f"{requests} requests, V lognormal median 16384 sigma 1.5 cap 2^22, degree 8, "
                    "50% ER / 50% R-MAT, seed 8",
        "capacity_graphs_per_s": capacity, "capacity_wall_s": cap_wall,
        "arrival_rate_per_s": rate, "load_fraction_of_capacity": load,
        "graphs_per_s": requests / wall, "p50_ms": p50, "p99_ms": p99, "mean_ms": float(lat.mean()),
        "hbm_budget_bytes": budget_gib << 30, "slots": slots,
        "completed": st.completed - st0.completed, "batches": st.batches - st0.batches,
        "solo": st.solo - st0.solo, "peak_admitted_bytes": st.peak_bytes,
        "warmup": f"the {slots} largest requests submitted {slots} times, then the capacity run",
    }


def bench_partitioned(steps=2):
    """-a d on the 8,388,608-vertex graph with 200 planted communities (mean
    degree 12, 2% inter-community arcs), partitioned by community: local
    SCCs per partition, then the merge (quotient-graph SCC)."""
    import detectingscc_amd as scc
    with scc.Context(0) as ctx:
        t0 = time.perf_counter()
        ctx.gen_communities_dev(1 << 23, 200, 12.0, 0.02, seed=55)   # also installs the partition
        g = ctx.graph
        log(f"generated the community graph on the device: V={g.V} E={g.E} in {time.perf_counter() - t0:.2f}s")
        counted = ctx.run(algo="d", sizes=True)
        ctx.run(algo="d", sizes=False)
        runs = [ctx.run(algo="d", sizes=False) for _ in range(steps)]
        phased = ctx.run(algo="d", sizes=False, phase_times=True)
    assert all(r.num_sccs == counted.num_sccs for r in runs)
    ms = sorted(r.ms_scc for r in runs)[len(runs) // 2]
    return {"workload": "8,388,608 V, 200 planted communities (seeded permutation), mean degree 12, 2% "
                        "inter-community arcs, seed 55; partition = community id (200 partitions)",
            "V": g.V, "E": g.E, "num_sccs": counted.num_sccs, "largest_scc": counted.largest,
            "ms": ms, "edges_per_s": g.E / (ms * 1e-3),
            "phase_ms": {k: round(v, 4) for k, v in phased.ms_phase.items() if v > 0}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat24", choices=sorted(WORKLOADS))
    ap.add_argument("--algo", default="g")
    ap.add_argument("--warp", type=int, default=1)
    ap.add_argument("--expand", type=int, default=4, help="0 auto, 1 vertex, 2 vwarp, 3 LB, 4 DO")
    ap.add_argument("--quick", action="store_true", help="the headline workload alone (no extra keys, no CPU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serial", action="store_true",
                    help="batch4096: skip the one-graph-per-launch baseline (for profiling the batched launch)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    import detectingscc_amd as scc
    if args.copy_calibrate:
        with scc.Context(0) as ctx:
            ms = ctx.copy_bench(args.copy_calibrate, 5)
        print(json.dumps({"copy_bytes": args.copy_calibrate, "ms_per_copy": ms,
                          "GBps_read_plus_write": 2 * args.copy_calibrate / ms / 1e6}), flush=True)
        return
    if args.workload in ("batch4096", "serve"):
        part = (bench_batched(steps=args.steps, serial_baseline=not args.no_serial) if args.workload == "batch4096"
                else bench_serving())
        print(json.dumps({
            "metric": "batched graphs/s and p99 vs one-graph-per-launch" if args.workload == "batch4096"
                      else "served graphs/s and p50/p99 latency under an HBM budget",
            "value": part["graphs_per_s"], "unit": "graphs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic", "config": {"workload": part.pop("workload"), "parallelism": "single GPU"},
            "detail": part}), flush=True)
        return

    if args.workload == "part200":
        part = bench_partitioned(steps=args.steps)
        print(json.dumps({
            "metric": "SCC edges/s of -a d (local SCCs per partition, then the merge)",
            "value": part["edges_per_s"], "unit": "edges/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
            "ms_per_step": part["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic",
            "config": {"workload": part.pop("workload"), "parallelism": "single GPU"}, "detail": part}), flush=True)
        return

    ctx = scc.Context(0)
    head = bench_single(ctx, args.workload, args.steps, args.warmup, args.algo, args.warp, args.expand)
    traffic, traffic_src = profiled_traffic(args.workload)
    fr_ms = head["search_ms"]
    achieved = head["search_GBps"] or 0.0
    out = {
        "metric": "SCC edges/s & %HBM peak; batched graphs/s and p99 vs one-graph-per-launch",
        "value": head["value"],
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": head["workload"], "V": head["V"], "E": head["E"], "algo": args.algo,
                   "warp": args.warp, "expand": args.expand, "num_sccs": head["num_sccs"],
                   "largest_scc": head["largest_scc"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "traffic_frac": traffic / (fr_ms * 1e6) / HBM_PEAK_GBS if traffic and fr_ms > 0 else None,
                     "kernel": "expand_do_kernel (first-pivot FW+BW search, direction-optimised)",
                     "search_ms": fr_ms,
                     "algorithmic_bytes": head["search_work_bytes"],
                     "work_units": head["search_work_units"],
                     "model_bytes": head["search_model_bytes"],
                     "model_frac": head["search_model_bytes"] / (fr_ms * 1e6) / HBM_PEAK_GBS if fr_ms > 0 else None,
                     "fw": head["fw"], "bw": head["bw"]},
        "phase_ms": head.get("phase_ms"),
        "phase_edges_per_s": head.get("phase_edges_per_s"),
        "device_ms_per_step": head["device_ms_per_step"],
        "histogram_device_ms": head["histogram_device_ms"],
        "labels_checked": head["labels_checked"],
    }
    if not args.no_cpu and not args.quick:
        g = ctx.download()   # the same graph, on the host
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(g, args.workload, threads)
        log(f"cpu baseline: {out['cpu_baseline']['ms_per_decomposition']:.0f} ms on {threads} threads, "
            f"Tarjan {out['cpu_baseline']['tarjan_single_thread_ms']:.0f} ms")
        del g
    else:
        out["cpu_baseline"] = None
    if not args.quick:
        for name, steps in (("rmat22", args.steps), ("grid4096", 3), ("er16k", args.steps)):
            if name == args.workload:
                continue
            r = bench_single(ctx, name, steps, 1)
            if name == "rmat22":
                r["matrix"] = bench_matrix(ctx, ctx.graph)
            out[name] = r
        ctx.close()
        out["batched"] = bench_batched()
        out["serving"] = bench_serving()
        out["partitioned"] = bench_partitioned()
    else:
        ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
