"""Benchmark: SCC decomposition throughput on one MI355X.

A step is one full SCC decomposition of a device-resident graph: CSR and
reverse CSR already in HBM when the timed region starts, canonical labels in
HBM when it ends.  Default workload: R-MAT scale 22 (4,194,304 vertices,
67,108,864 arcs, a=0.57 b=0.19 c=0.19, seeded permutation).

Printed as ONE JSON line:
  value        SCC edges/s = E * steps / wall time of the timed steps
  roofline     the dominant kernel (frontier expansion of the first-pivot
               FW+BW search): algorithmic bytes 12|R| + 8|E_R| per direction
               over the device time of that search (HIP events on the
               library's stream), against 8 TB/s HBM.  Bottom-up passes stop
               each vertex at its first visited parent and read far fewer
               arcs, so `frac` can pass 1; `traffic` is the search's HBM
               bytes measured with rocprof counters and `traffic_frac` those
               bytes over the same time against 8 TB/s
  cpu_baseline the CPU restatement of vHong (oracle port), all host cores,
               on a bounded sample (R-MAT scale 20)
  batched      the 4096-graph mix (V log-uniform in [2^8, 2^16], mean degree
               uniform in [2, 16], 40% ER / 40% R-MAT / 20% planted cycles):
               autotuned persistent launches vs one graph per launch through
               the single-graph engine, graphs/s and per-graph p50/p99
  serving      Poisson arrivals of lognormal requests (median 2^14 V, sigma
               1.5, capped at 2^22, mean degree 8, ER and R-MAT) into the
               continuous-batching server under an HBM budget: graphs/s and
               p50/p99 latency (submit -> labels on the host)
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
    # name: (description, builder)
    "rmat22": "R-MAT scale 22, 4,194,304 V, 67,108,864 arcs, a=.57 b=.19 c=.19, seed 22",
    "rmat24": "R-MAT scale 24, 16,777,216 V, 268,435,456 arcs, a=.57 b=.19 c=.19, seed 24",
    "er16k": "ER 16,384 V, 131,072 draws + 64 planted cycles (2..256), seed 1",
    "grid4096": "grid 4096x4096, arc probability 0.7, seed 4096",
    # measured alone (one JSON line each, value = graphs/s)
    "batch4096": "the 4096-graph mix, batched persistent form vs one graph per launch",
    "serve": "Poisson request stream into the continuous-batching server",
}


def build_graph(name):
    import detectingscc_amd as scc
    if name == "rmat22":
        return scc.gen_rmat(22, seed=22)
    if name == "rmat24":
        return scc.gen_rmat(24, seed=24)
    if name == "er16k":
        return scc.gen_er(16384, 131072, cycles=64, cmin=2, cmax=256, seed=1)
    if name == "grid4096":
        return scc.gen_grid(4096, 0.7, seed=4096)
    raise SystemExit(f"unknown workload {name}")


def reach_bytes(nv, ne):
    return 12 * nv + 8 * ne


def profiled_traffic(workload):
    """HBM bytes of one first-pivot FW+BW search measured with rocprofv3
    (FETCH_SIZE + WRITE_SIZE, calibrated on a plain copy), if recorded."""
    p = os.path.join(ROOT, "profiles", f"reach_traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_search")


def cpu_baseline(sample_scale: int, threads: int):
    import detectingscc_amd as scc
    from oracle import pyoracle as orc
    g = scc.gen_rmat(sample_scale, seed=22)
    orc.vhong(g, True, True, threads)          # warm-up
    runs = []
    t_end = time.perf_counter() + 10.0
    while time.perf_counter() < t_end or len(runs) < 3:
        t = time.perf_counter()
        orc.vhong(g, True, True, threads)
        runs.append(time.perf_counter() - t)
        if len(runs) >= 50:
            break
    runs.sort()
    med = runs[len(runs) // 2]
    return {"value": g.E / med, "unit": "edges/s", "cores": threads or orc.max_threads(),
            "kind": "port",
            "sample": f"R-MAT scale {sample_scale} ({g.V} V, {g.E} arcs), vHong phase order, "
                      f"median of {len(runs)} runs"}


def bench_batched(graphs=4096, steps=3, sample=1024, serial_baseline=True):
    import detectingscc_amd as scc
    t0 = time.perf_counter()
    b = scc.gen_batch_mix(graphs, 8, 16, 2.0, 16.0, seed=4096)
    log(f"built the {graphs}-graph mix: {int(b.V.sum())} V, {int(b.E.sum())} arcs "
        f"in {time.perf_counter() - t0:.1f}s")
    with scc.Context(0) as ctx:
        ctx.batch_upload(b)
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
        "vertices": int(b.V.sum()), "arcs": int(b.E.sum()),
        "graphs_per_s": best.graphs_per_s, "ms": best.ms, "p50_ms": best.p50_ms, "p99_ms": best.p99_ms,
        "launches": best.launches, "steps": steps,
        "serial_graphs_per_s": serial.graphs_per_s if serial else None,
        "serial_ms": serial.ms if serial else None,
        "serial_p99_ms": serial.p99_ms if serial else None,
        "speedup_vs_one_graph_per_launch": best.graphs_per_s / serial.graphs_per_s if serial else None,
        "labels_identical": same,
        "autotune": report.strip().splitlines(),
    }


def bench_serving(requests=400, rate=1000.0, budget_gib=32, slots=3):
    import detectingscc_amd as scc
    t0 = time.perf_counter()
    b = scc.gen_serving_mix(requests, 16384, 1.5, 1 << 22, 8.0, seed=8)
    graphs = [b.graph(i) for i in range(requests)]
    log(f"built {requests} requests: {int(b.V.sum())} V, {int(b.E.sum())} arcs, largest {int(b.V.max())} V "
        f"in {time.perf_counter() - t0:.1f}s")
    gaps = np.random.default_rng(1).exponential(1.0 / rate, size=requests)
    with scc.Context(0) as c:
        c.server_start(budget_bytes=budget_gib << 30, max_batch=256, slots=slots, solo_vertices=1 << 16,
                       lanes=16, block=1024)
        # warm-up (untimed): the largest requests, a few times, so every
        # slot's grow-only device buffers reach their size before the run
        big = sorted(range(requests), key=lambda i: -graphs[i].E)[:slots]
        for _ in range(slots):
            warm_ids = [c.server_submit(graphs[i]) for i in big]
            c.server_wait_all()
            for rid in warm_ids:
                c.server_poll(rid)
        warm = c.server_stats()
        t = time.perf_counter()
        due = t
        ids = []
        for g, gap in zip(graphs, gaps):
            due += gap
            while time.perf_counter() < due:
                pass
            ids.append(c.server_submit(g))
        c.server_wait_all()
        wall = time.perf_counter() - t
        st = c.server_stats()
        lat = np.array([c.server_poll(rid)[2] for rid in ids])
        c.server_stop()
    p50, p99 = (float(x) for x in np.percentile(lat, [50, 99]))
    log(f"served {len(ids)} requests in {wall:.2f}s: p50 {p50:.2f} ms, p99 {p99:.2f} ms")
    return {
        "workload": f"{requests} requests, Poisson {rate:g}/s, V lognormal median 16384 sigma 1.5 cap 2^22, "
                    "degree 8, 50% ER / 50% R-MAT, seed 8",
        "arrival_rate_per_s": rate, "hbm_budget_bytes": budget_gib << 30, "slots": slots,
        "graphs_per_s": len(ids) / wall, "p50_ms": p50, "p99_ms": p99, "mean_ms": float(lat.mean()),
        "completed": st.completed - warm.completed, "batches": st.batches - warm.batches,
        "solo": st.solo - warm.solo, "peak_admitted_bytes": st.peak_bytes,
        "warmup": f"the {slots} largest requests submitted {slots} times before the timed arrivals",
    }


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22", choices=sorted(WORKLOADS))
    ap.add_argument("--algo", default="g")
    ap.add_argument("--warp", type=int, default=1)
    ap.add_argument("--expand", type=int, default=4, help="0 auto, 1 vertex, 2 vwarp, 3 LB, 4 DO")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-scale", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-mix measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serve-rate", type=float, default=1000.0)
    ap.add_argument("--no-serial", action="store_true",
                    help="batched mix: skip the one-graph-per-launch baseline (for profiling the batched launch)")
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
                else bench_serving(rate=args.serve_rate))
        print(json.dumps({
            "metric": "batched graphs/s and p99 vs one-graph-per-launch" if args.workload == "batch4096"
                      else "served graphs/s and p50/p99 latency under an HBM budget",
            "value": part["graphs_per_s"], "unit": "graphs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic", "config": {"workload": part.pop("workload"), "parallelism": "single GPU"},
            "detail": part}), flush=True)
        return
    t0 = time.perf_counter()
    g = build_graph(args.workload)
    print(f"[bench] built {args.workload}: V={g.V} E={g.E} in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    ctx = scc.Context(0)
    ctx.upload(g)
    kw = dict(algo=args.algo, warp=args.warp, expand=args.expand, sizes=False)
    # one untimed decomposition counts |R| and |E_R| of the first search (the
    # searches are deterministic, so the timed runs skip those sweeps)
    counted = ctx.run(**dict(kw, sizes=True))
    for i in range(args.warmup):
        st = ctx.run(**kw)
        print(f"[bench] warmup {i}: {st.ms_scc:.3f} ms device, {st.num_sccs} SCCs", file=sys.stderr, flush=True)
    stats = []
    t = time.perf_counter()
    for _ in range(args.steps):
        stats.append(ctx.run(**kw))     # synchronous: returns after the last kernel
    wall = time.perf_counter() - t
    ms_step = 1e3 * wall / args.steps
    st = stats[-1]
    # dominant kernel: first-pivot FW+BW frontier expansion
    fr = [s.ms_first_reach for s in stats]
    fr_ms = sorted(fr)[len(fr) // 2]
    alg = (reach_bytes(counted.first_fw_vertices, counted.first_fw_arcs) +
           reach_bytes(counted.first_bw_vertices, counted.first_bw_arcs))
    achieved = alg / (fr_ms * 1e-3) / 1e9 if fr_ms > 0 else 0.0
    traffic = profiled_traffic(args.workload)
    traffic_frac = traffic / (fr_ms * 1e6) / HBM_PEAK_GBS if traffic and fr_ms > 0 else None
    # the phase breakdown comes from untimed decompositions that record a
    # device timestamp at every phase boundary (each costs a few us of idle
    # GPU, so the timed ones record only their start, end and first search)
    phased = [ctx.run(**dict(kw, phase_times=True)) for _ in range(max(1, min(args.steps, 5)))]
    phase_ms = {k: round(sum(s.ms_phase[k] for s in phased) / len(phased), 4) for k in scc.PHASES}
    out = {
        "metric": "SCC edges/s & %HBM peak; batched graphs/s and p99 vs one-graph-per-launch",
        "value": g.E * args.steps / wall,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload], "V": g.V, "E": g.E, "algo": args.algo,
                   "warp": args.warp, "expand": args.expand, "num_sccs": st.num_sccs,
                   "largest_scc": counted.largest, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_frac": traffic_frac,
                     "kernel": "expand (first-pivot FW+BW)", "search_ms": fr_ms,
                     "algorithmic_bytes": alg,
                     "fw": [counted.first_fw_vertices, counted.first_fw_arcs],
                     "bw": [counted.first_bw_vertices, counted.first_bw_arcs]},
        "phase_ms": phase_ms,
        "device_ms_per_step": sum(s.ms_scc for s in stats) / len(stats),
    }
    ctx.close()
    del g
    out["batched"] = None if args.no_batch else bench_batched()
    out["serving"] = None if args.no_serve else bench_serving(rate=args.serve_rate)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_scale, 0)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
