#!/usr/bin/env python3
"""Benchmark: RTRExtractor decomposition (count -> clean -> extract) on MI355X.

One step = one full decomposition of the workload graph, whose device CSR was
built before timing: weighted triangle supports, cascading clean to
convergence, greedy extraction, and the results -- cluster id of every vertex
and the packed survivor bitmap -- in host memory (page-locked buffers the
device writes directly).  Prints ONE JSON line; `value` is that rate for the
medium planted-community workload (C2).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--eps 0.3]
                    [--workload c2|c1|c3|rmat20] [--no-cpu] [--c3-steps S]
                    [--batch-graphs B] [--serve-seconds T] [--serve-capacity R1,R2]

Beside the headline: the same steps with the results left in HBM, with the
upload and device CSR build, the skewed large R-MAT scale-22 graph (C3), the
4,096-graph batch against the plain path (C4, with its first call cold), and a
60 s Poisson serving trace with p50/p99 latency overall and per request kind
(C5).

Roofline: the count phase (its kernels' HIP-event time on the library's
stream) priced with the algorithmic bytes B = 8*S_min + 28*m + 16*n, S_min =
sum over edges of min(d_x, d_y) (4-byte id streamed + 4-byte probe per element
of the shorter list, CSR read once, two FP64 supports per edge, one per
vertex).  Peak: MI355X HBM3E 8 TB/s.  `traffic` is the HBM bytes the count
kernels moved per phase (FETCH_SIZE x 2 + WRITE_SIZE, from rocprofv3; every
read of these kernels is a 128-byte request, which FETCH_SIZE counts as 64 B on
gfx950), read from profiles/ when present.

CPU baseline: the reference's own sequential code (oracle/_ref/ref_dump, built
from the reference sources) on the same graph, 1 core, when that binary is
present; otherwise the in-repo C++ restatement (oracle/liboracle.so).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
T_START = time.perf_counter()


def progress(msg: str) -> None:
    """A progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench.py [{time.perf_counter() - T_START:6.1f} s] {msg}", file=sys.stderr, flush=True)


def workload(name: str):
    from rtrex_cuda_amd import graphs as G
    if name == "c1":
        return ("C1 Chung-Lu n=10,000 m=100,000 gamma=2.5 dmax=500 seed=1", *G.config1())
    if name == "c3":
        return ("C3 R-MAT scale 22 edge factor 16 (0.57,0.19,0.19) seed=11", *G.config3())
    if name == "rmat20":
        return ("R-MAT scale 20 edge factor 16 (0.57,0.19,0.19), the serving pool's scale-20 "
                "graph", *G.serve_graph("rmat", 2))
    return ("C2 planted-community n=335,000 (5,000 communities of 10-80, p=0.3) + "
            "Chung-Lu background gamma=2.2, seed=7", *G.config2())


def algorithmic_count_bytes(n, u, v) -> tuple[int, int]:
    deg = np.bincount(np.concatenate([u, v]), minlength=n).astype(np.int64)
    s_min = int(np.minimum(deg[u], deg[v]).sum())
    m = len(u)
    return 8 * s_min + 28 * m + 16 * n, s_min


def batched_vs_plain(eps: float, count: int, steps: int) -> dict:
    """C4: `count` planted-community graphs resident at once as one batch vs
    the same graphs through the plain path one after another (each graph
    already resident), same eps."""
    import rtrex_cuda_amd as R
    from rtrex_cuda_amd import graphs as G
    t0 = time.perf_counter()
    graphs = G.config4(count)
    gen_s = time.perf_counter() - t0
    m = int(sum(len(g[1]) for g in graphs))
    n = int(sum(g[0] for g in graphs))
    B_count = sum(algorithmic_count_bytes(*g)[0] for g in graphs)
    t0 = time.perf_counter()
    b = R.Batch(graphs)
    build_s = time.perf_counter() - t0
    # The first call on the fresh batch: count plans, triangle-list sizing,
    # and the discovery and exact re-run of undecided regions (later calls at
    # the same eps re-run the remembered regions directly).
    t0 = time.perf_counter()
    b.decompose(eps)
    cold_ms = (time.perf_counter() - t0) * 1e3
    params, tbest, tstart = b.autotune(repeats=2)
    cmb = R.host_empty(n, np.int64)  # page-locked: the device writes the ids
    b.decompose(eps, out=cmb)
    tms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        maps, nc = b.decompose(eps, out=cmb)
        tms.append(b.timings())
    bt = (time.perf_counter() - t0) / steps
    audit = b.audit()
    b.close()
    # Plain path on every 8th graph: each graph resident (built, warmed up),
    # then one timed decomposition; throughput = sampled edges / summed time.
    stride = 8
    pt = 0.0
    pm = 0
    same = 0
    sample = list(range(0, count, stride))
    for i in sample:
        with R.DeviceGraph.from_edges(*graphs[i]) as p:
            po = R.host_empty(graphs[i][0], np.int64)
            p.decompose(eps, out=po)
            t0 = time.perf_counter()
            c2, _ = p.decompose(eps, out=po)
            pt += time.perf_counter() - t0
        pm += len(graphs[i][1])
        same += int(np.array_equal(c2, maps[i]))
    cnt_ms = float(np.mean([t["count_ms"] for t in tms]))
    ach = B_count / (cnt_ms / 1e3) / 1e9
    return {"workload": f"C4 {count} planted-community graphs, n log-uniform [50, 20000], "
                        "communities 5-40 p=0.4, background mean degree 6, seed 23",
            "graphs": count, "n": n, "m": m, "generation_s": round(gen_s, 1),
            "batched_value": round(m / bt / 1e6, 3), "batched_ms": round(bt * 1e3, 3),
            "upload_and_build_s": round(build_s, 3),
            "first_call_cold_ms": round(cold_ms, 3),
            "plain_value": round(pm / pt / 1e6, 3),
            "plain_sample": f"{len(sample)} of {count} graphs (every {stride}th), "
                            f"{pm} edges in {pt:.3f} s",
            "speedup": round((m / bt) / (pm / pt), 2),
            "per_graph_identical": f"{same}/{len(sample)} sampled graphs equal alone vs in batch",
            "autotune": {"params": params, "count_ms_default": round(tstart, 3),
                         "count_ms_tuned": round(tbest, 3)},
            "phases_ms": {k: round(float(np.mean([t[k] for t in tms])), 3)
                          for k in ("count_ms", "clean_ms", "extract_ms")},
            "count_roofline": {"achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(ach / HBM_PEAK_GBS, 4),
                               "algorithmic_bytes": int(B_count)},
            "audit": audit}


def copy_calibration(nbytes: int, iters: int) -> int:
    import ctypes as C

    from rtrex_cuda_amd import build as B
    L = C.CDLL(B.BENCH_LIB)
    L.rtrex_bench_copy.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    L.rtrex_bench_copy.restype = C.c_int
    ms = C.c_double()
    if L.rtrex_bench_copy(0, nbytes, iters, C.byref(ms)):
        print("bench.py: calibration copy failed", file=sys.stderr)
        return 1
    print(json.dumps({"copy_bytes": nbytes, "launches": iters + 1, "ms_per_launch": ms.value,
                      "GB/s (read+write)": 2 * nbytes / (ms.value / 1e3) / 1e9}), flush=True)
    return 0


def run_trace(R, pool, trace, eps: float, workers: int, budget_gb: float, note: bool = True):
    """Submits a (time, kind, index) trace to a fresh server at its scheduled
    times; returns per request (clusters, job info, map hash) and the clock
    origin of the schedule."""
    import queue
    import threading

    graph = lambda kind, i: pool[kind][i]  # noqa: E731
    done = [None] * len(trace)
    with R.Server(workers=workers, hbm_budget_bytes=int(budget_gb * (1 << 30))) as srv:
        for kind in ("cl", "planted", "rmat"):  # warm the code paths once
            srv.wait(srv.submit(*graph(kind, 0), eps))
        q: "queue.Queue" = queue.Queue()

        def waiter():
            while True:
                item = q.get()
                if item is None:
                    return
                i, j = item
                cm, k, info = srv.wait(j)
                done[i] = (k, info, hashlib.sha256(cm.tobytes()).digest())

        th = threading.Thread(target=waiter)
        th.start()
        start = srv.clock() + 0.05
        next_note = 10.0
        for i, (t, kind, idx) in enumerate(trace):
            if note and t >= next_note:
                progress(f"c5: {t:.0f} s of the trace submitted")
                next_note += 10.0
            while srv.clock() < start + t:
                time.sleep(min(0.0005, max(0.0, start + t - srv.clock())))
            q.put((i, srv.submit(*graph(kind, idx), eps)))
        q.put(None)
        th.join()
    return done, start


def serving_leg(eps: float, seconds: float, rate: float, workers: int, budget_gb: float,
                capacity_rates=(), capacity_seconds: float = 10.0) -> dict:
    """C5: Poisson request stream through the server (continuous batching,
    HBM-footprint admission) vs the same requests one at a time through the
    plain path.  Latency = result completion - scheduled arrival.  Capacity:
    shorter traces of the same mix at higher rates; the highest rate whose p99
    stays under a second with every request served at the offered rate."""
    import rtrex_cuda_amd as R
    from rtrex_cuda_amd import graphs as G
    sizes = (64, 32, 3)
    t0 = time.perf_counter()
    pool = G.serve_pool(*sizes)
    gen_s = time.perf_counter() - t0
    trace = G.serve_trace(seconds, rate, pool_sizes=sizes)
    graph = lambda kind, i: pool[kind][i]  # noqa: E731
    done, start = run_trace(R, pool, trace, eps, workers, budget_gb)
    lat = np.array([done[i][1]["end_s"] - (start + t) for i, (t, _, _) in enumerate(trace)])
    kinds = np.array([k for _, k, _ in trace])
    by_kind = {}
    for kind in ("cl", "planted", "rmat"):
        lk = lat[kinds == kind]
        if len(lk):
            by_kind[kind] = {"requests": int(len(lk)),
                             "p50": round(float(np.percentile(lk, 50)) * 1e3, 2),
                             "p99": round(float(np.percentile(lk, 99)) * 1e3, 2)}
    edges = np.array([len(graph(k, i)[1]) for _, k, i in trace], np.float64)
    makespan = max(d[1]["end_s"] for d in done) - start
    batched = np.array([d[1]["batch_graphs"] for d in done])
    # Capacity sweep: the same mix at higher arrival rates.
    sweep = []
    for r in capacity_rates:
        tr = G.serve_trace(capacity_seconds, r, pool_sizes=sizes)
        dn, st0 = run_trace(R, pool, tr, eps, workers, budget_gb, note=False)
        lt = np.array([dn[i][1]["end_s"] - (st0 + t) for i, (t, _, _) in enumerate(tr)])
        ed = float(sum(len(graph(k, i)[1]) for _, k, i in tr))
        span = max(d[1]["end_s"] for d in dn) - st0
        row = {"rate": r, "requests": len(tr), "p50_ms": round(float(np.percentile(lt, 50)) * 1e3, 2),
               "p99_ms": round(float(np.percentile(lt, 99)) * 1e3, 2),
               "offered_Medges_per_s": round(ed / capacity_seconds / 1e6, 2),
               "served_Medges_per_s": round(ed / span / 1e6, 2)}
        sweep.append(row)
        progress(f"c5: {r:g} req/s p99 {row['p99_ms']:.0f} ms")
    main_ok = (float(np.percentile(lat, 99)) < 1.0 and
               edges.sum() / makespan >= 0.97 * edges.sum() / seconds)
    ok = ([rate] if main_ok else []) + [
        row["rate"] for row in sweep
        if row["p99_ms"] < 1000.0 and row["served_Medges_per_s"] >= 0.97 * row["offered_Medges_per_s"]]
    # Capacity: the first requests of the same trace submitted at once (no
    # arrival process); throughput = their edges / (last completion - first
    # submission).
    cap_n = min(len(trace), 2000)
    with R.Server(workers=workers, hbm_budget_bytes=int(budget_gb * (1 << 30))) as srv:
        for kind in ("cl", "planted", "rmat"):
            srv.wait(srv.submit(*graph(kind, 0), eps))
        t_sub = srv.clock()
        ids = [srv.submit(*graph(kind, idx), eps) for _, kind, idx in trace[:cap_n]]
        cap_end = max(srv.wait(j)[2]["end_s"] for j in ids)
    cap_edges = float(sum(len(graph(k, i)[1]) for _, k, i in trace[:cap_n]))
    cap_rate = cap_edges / (cap_end - t_sub) / 1e6
    progress(f"c5: saturated {cap_rate:.1f} M edges/s")
    # Serial plain runs on a sample (every 10th request), inputs in host memory
    # as the server gets them: upload + build + decomposition per request.
    sample = list(range(0, len(trace), 10))
    st, sm, same = 0.0, 0.0, 0
    for i in sample:
        _, kind, idx = trace[i]
        n, u, v = graph(kind, idx)
        t1 = time.perf_counter()
        with R.DeviceGraph.from_edges(n, u, v) as g:
            cm, k = g.decompose(eps)
        st += time.perf_counter() - t1
        sm += len(u)
        same += int(k == done[i][0] and hashlib.sha256(cm.tobytes()).digest() == done[i][2])
    return {"workload": f"C5 Poisson {rate:g} req/s for {seconds:g} s, mix 70% Chung-Lu "
                        "m in [1k,50k], 25% planted m in [50k,2M], 5% R-MAT scale 18-20; "
                        f"graphs from a seeded pool of {sum(sizes)} (seed 42)",
            "requests": len(trace), "pool_generation_s": round(gen_s, 1),
            "offered_Medges_per_s": round(edges.sum() / seconds / 1e6, 2),
            "served_Medges_per_s": round(edges.sum() / makespan / 1e6, 2),
            "makespan_s": round(makespan, 3),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                           "p99": round(float(np.percentile(lat, 99)) * 1e3, 2),
                           "max": round(float(lat.max()) * 1e3, 2), "by_kind": by_kind},
            "mean_graphs_per_launch": round(float(batched.mean()), 2),
            "workers": workers, "hbm_budget_GB": budget_gb,
            "serial_plain_Medges_per_s": round(sm / st / 1e6, 2),
            "serial_plain_sample": f"{len(sample)} requests (every 10th), {st:.2f} s",
            "served_vs_serial": round((edges.sum() / makespan) / (sm / st), 2),
            "capacity": {"seconds_per_rate": capacity_seconds, "rates": sweep,
                         "highest_rate_p99_under_1s": max(ok) if ok else None,
                         "note": "traces of the same mix at higher Poisson rates; a rate "
                                 "counts when p99 < 1 s and served >= 97 % of offered"},
            "saturated": {"requests": cap_n, "Medges_per_s": round(cap_rate, 2),
                          "vs_serial": round(cap_rate / (sm / st / 1e6), 2),
                          "note": "the trace's first requests submitted at once"},
            "same_cluster_map_as_alone": f"{same}/{len(sample)} sampled requests"}


def plain_leg(R, n, u, v, eps, steps, warmup, tune, upload) -> dict:
    """One graph through the plain path: `steps` timed decompositions with the
    results in host memory, the same with the results left in HBM, and
    (upload) with the edge list uploaded and the device CSR built each step."""
    m = len(u)
    g = R.DeviceGraph.from_edges(n, u, v)  # resident in HBM before timing
    cm = R.host_empty(n, np.int64)  # page-locked: the device writes the ids
    words = R.host_empty((m + 31) // 32, np.uint32)
    for _ in range(1 + warmup):  # the first call builds the count plan and lists
        g.decompose(eps, out=cm)
        g.survivor_words(out=words)
    phases = []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, k = g.decompose(eps, out=cm)  # cluster ids in host memory
        g.survivor_words(out=words)     # survivor bitmap in host memory
        phases.append(g.timings())
    ms = (time.perf_counter() - t0) * 1e3 / steps
    audit = g.audit()
    survivors = int(np.unpackbits(words.view(np.uint8))[:m].sum())
    # The same steps with both results left in device memory.
    nr = max(1, steps // 2)
    t0 = time.perf_counter()
    for _ in range(nr):
        g.decompose(eps, host_map=False)
    ms_res = (time.perf_counter() - t0) * 1e3 / nr
    tuned = {}
    if tune:
        params, best, start = g.autotune(repeats=2)
        tuned = {"params": params, "count_ms_default": round(start, 3),
                 "count_ms_tuned": round(best, 3)}
    g.close()
    out = {"ms_per_step": ms, "clusters": k, "survivors": survivors, "audit": audit,
           "autotune": tuned,
           "phases_ms": {key: round(float(np.mean([p[key] for p in phases])), 3)
                         for key in ("count_ms", "clean_ms", "extract_ms",
                                     "count_edge_kernel_ms", "count_vertex_kernel_ms")},
           "results_resident": {"value": round(m / (ms_res / 1e3) / 1e6, 3),
                                "ms_per_step": round(ms_res, 3),
                                "note": "the same steps, cluster ids and survivor bits left "
                                        "in HBM"}}
    if upload:
        nb_steps = max(1, steps // 3)
        t0 = time.perf_counter()
        for _ in range(nb_steps):
            with R.DeviceGraph.from_edges(n, u, v) as g2:
                g2.decompose(eps, out=cm)
                g2.survivor_words(out=words)
        ms_build = (time.perf_counter() - t0) * 1e3 / nb_steps
        out["with_upload_and_build"] = {
            "value": round(m / (ms_build / 1e3) / 1e6, 3), "ms_per_step": round(ms_build, 3),
            "note": "host edge list -> device CSR build -> decomposition -> results on the host"}
    return out


def c3_leg(R, eps: float, steps: int) -> dict:
    """C3: the skewed large R-MAT scale-22 graph through the plain path, and
    its count phase against the roofline (S_min computed exactly)."""
    desc, n, u, v = workload("c3")
    progress(f"c3: generated, m={len(u)}")
    B, s_min = algorithmic_count_bytes(n, u, v)
    t0 = time.perf_counter()
    r = plain_leg(R, n, u, v, eps, steps, 0, False, upload=False)
    total_s = time.perf_counter() - t0
    m = len(u)
    cnt_ms = r["phases_ms"]["count_ms"]
    ach = B / (cnt_ms / 1e3) / 1e9
    return {"workload": desc, "n": n, "m": m, "eps": eps, "steps": steps,
            "value": round(m / (r["ms_per_step"] / 1e3) / 1e6, 3), "unit": "M edges/s",
            "ms_per_step": round(r["ms_per_step"], 3), "phases_ms": r["phases_ms"],
            "clusters": r["clusters"], "survivors": r["survivors"], "audit": r["audit"],
            "count_roofline": {"achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes": B,
                               "s_min": s_min},
            "leg_wall_s": round(total_s, 1),
            "note": "graph resident (its build and first decomposition untimed), results in "
                    "host memory"}


def host_cpu() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_baseline(n, u, v, eps) -> dict:
    import oracle as O
    if O.ref_available():
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "g.txt")
            with open(p, "w") as f:
                f.write(f"{n} {len(u)}\n")
                f.write("\n".join(f"{a} {b}" for a, b in zip(u.tolist(), v.tolist())))
                f.write("\n")
            r = O.ref_dump(p, eps, extract=True)
        t = r["times"]["full_pipeline"]  # disjointExtract: count + clean + extract
        return {"value": round(len(u) / t / 1e6, 4), "unit": "M edges/s", "cores": 1,
                "kind": "reference", "host": host_cpu(),
                "sample": f"full workload once through the reference's disjointExtract "
                          f"(RTRex_Sequential, mode s), {t:.2f} s; file load excluded"}
    off, nb = O.csr_from_undirected(n, u, v)
    t0 = time.perf_counter()
    O.decompose(off, nb, eps, with_clusters=False)
    t = time.perf_counter() - t0
    return {"value": round(len(u) / t / 1e6, 4), "unit": "M edges/s", "cores": 1,
            "kind": "port", "host": host_cpu(),
            "sample": f"full workload once through the in-repo C++ restatement, {t:.2f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--eps", type=float, default=0.3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "rmat20"],
                    help="c2 is the headline; c1, c3 and rmat20 for profiling")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the count-plan autotuner (default parameters)")
    ap.add_argument("--batch-graphs", type=int, default=4096,
                    help="C4 batched-vs-plain leg (0 skips it)")
    ap.add_argument("--c3-steps", type=int, default=2,
                    help="C3 leg (R-MAT scale 22): timed decompositions (0 skips it)")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="C5 serving leg: length of the request trace (0 skips it)")
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-capacity", default="300,360",
                    help="C5 capacity sweep: comma-separated Poisson rates (req/s) run for 10 s "
                         "each after the main trace ('' skips it)")
    ap.add_argument("--serve-budget-gb", type=float, default=128.0,
                    help="serving pool size (HBM footprint admission), GiB of the 288 GB")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the calibration copy kernel over BYTES and print "
                         "its time (scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, max(1, args.steps))
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one MI355X only; runs on several GPUs are not supported "
              "(each graph or job fits one GPU; a larger graph would have its edges partitioned "
              "across GPUs, with the clean's support decrements exchanged once per peel round, "
              "which is not implemented)",
              file=sys.stderr)
        return 2

    import rtrex_cuda_amd as R
    desc, n, u, v = workload(args.workload)
    m = len(u)
    B_count, s_min = algorithmic_count_bytes(n, u, v)
    progress(f"{args.workload}: n={n} m={m}")
    plain = plain_leg(R, n, u, v, args.eps, args.steps, args.warmup, not args.no_tune,
                      upload=True)
    progress(f"{args.workload}: {plain['ms_per_step']:.3f} ms/step")
    ms = plain["ms_per_step"]
    value = m / (ms / 1e3) / 1e6
    cnt_ms = plain["phases_ms"]["count_ms"]
    achieved = B_count / (cnt_ms / 1e3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"count_{args.workload}_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_count_phase")

    out = {
        "metric": "M edges/s count+clean+extract (plain; graph resident in HBM, cluster ids "
                  "and survivor bitmap in host memory)",
        "value": round(value, 3),
        "unit": "M edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator; float32 edge weights generated but unused by "
                "the path, as in the reference)",
        "config": {"workload": desc, "n": n, "m": m, "eps": args.eps,
                   "clusters": plain["clusters"], "survivors": plain["survivors"]},
        "phases_ms": plain["phases_ms"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": "count phase (k_count_small, "
                     "k_count_hub_stream, k_count_hub, k_count_hub_large, k_count_hub_huge, "
                     "k_count_edges, k_sum_tcnt)",
                     "algorithmic_bytes": B_count, "s_min": s_min},
        "audit": plain["audit"],
        "autotune": plain["autotune"],
        "results_resident": plain["results_resident"],
        "with_upload_and_build": plain["with_upload_and_build"],
    }
    if args.c3_steps > 0 and args.workload != "c3":
        out["skewed_large"] = c3_leg(R, args.eps, args.c3_steps)
        progress(f"c3: {out['skewed_large']['ms_per_step']:.1f} ms/step")
    if args.batch_graphs > 0:
        out["batched"] = batched_vs_plain(args.eps, args.batch_graphs, max(1, args.steps // 3))
        progress(f"c4: {out['batched']['batched_ms']:.1f} ms/step batched")
    if args.serve_seconds > 0:
        rates = tuple(float(x) for x in args.serve_capacity.split(",") if x.strip())
        out["serving"] = serving_leg(args.eps, args.serve_seconds, 200.0, args.serve_workers,
                                     args.serve_budget_gb, rates)
        progress("c5: done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, u, v, args.eps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
