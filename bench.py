#!/usr/bin/env python3
"""bench.py -- HipGrid FirmCore decomposition on one MI355X.

One step = one full decomposition (core index for every lmd = 1..L, cell
lengths) of the workload graph, already resident in HBM, through the C ABI
(hg_decompose).  Prints ONE JSON line on stdout.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1]
                  [--cpu-baseline auto|reference|port|off]

Metric: GTEPS = L * D / t, D = directed CSR entries over all layers (every
lmd pass traverses every entry once).  Roofline: HBM, algorithmic bytes
B_alg = L * (12 D + 8 L (n+1) + 4 n L + 8 n) per decomposition.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FirmCore decomp time + GTEPS (% HBM roofline); batched graphs/s, p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(n: int, L: int, D: int) -> int:
    """Per decomposition: L passes of 4D ids + 8D degree atomics + 8L(n+1)
    per-layer offsets + 4nL degree init + 8n index/alive state."""
    return L * (12 * D + 8 * L * (n + 1) + 4 * n * L + 8 * n)


def measured_traffic(name: str, key: str = "traffic_bytes_per_decomposition"):
    """HBM bytes per launch from the committed rocprofv3 counter summary
    profiles/<name>.md (FETCH_SIZE and WRITE_SIZE runs, calibrated against
    a 1 GiB copy), or None when no summary exists."""
    path = os.path.join(ROOT, "profiles", f"{name}.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith(key + ":"):
            return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def host_info() -> dict:
    """The host the CPU baseline runs on: hardware threads and CPU model."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def run_reference_pargrid(g, threads: int, method: str = "ParGrid", extra=()) -> dict:
    """Runs the reference's own driver (oracle/_ref/firmcore_ref, built from
    its sources at -O3) as `-m <method> -num_thread threads` on g written in
    the reference's text format; returns its printed elapsed time (its own
    omp_get_wtime bracket, main.cpp:183-245 and 402-424) and its peak RSS
    (VmHWM sampled while it runs, MB)."""
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "firmcore_ref")
    tmp = tempfile.mkdtemp(prefix="hgcpu_")
    try:
        ds = os.path.join(tmp, "dataset", "bench")
        os.makedirs(ds)
        os.makedirs(os.path.join(tmp, "run"))
        g.write_text(ds)
        # The peak RSS the reference prints comes from getrusage, which keeps
        # the high-water mark of the process it was forked from (this one,
        # holding the benchmark graphs) across exec; its own is VmHWM of its
        # address space, sampled while it runs.
        out_path = os.path.join(tmp, "ref.out")
        fout = open(out_path, "w")  # a file, not a pipe: lpe prints a line per cell
        proc = subprocess.Popen([ref_bin, "-d", "bench", "-m", method, "-num_thread", str(threads), *extra],
                                cwd=os.path.join(tmp, "run"), stdout=fout, stderr=subprocess.STDOUT)
        hwm_kb = 0
        t_end = time.time() + 900
        while proc.poll() is None and time.time() < t_end:
            try:
                for line in open(f"/proc/{proc.pid}/status"):
                    if line.startswith("VmHWM:"):
                        hwm_kb = max(hwm_kb, int(line.split()[1]))
            except OSError:
                pass
            time.sleep(0.02)
        if proc.poll() is None:
            proc.kill()
        proc.wait()
        fout.close()
        stdout, stderr = open(out_path).read(), ""
        secs = None
        rss = hwm_kb / 1024.0 if hwm_kb else None
        marker = {"ParGrid": "Path left Core Parallel Sync Elapsed time", "lpe": "Serial Elapsed time",
                  "kpe": "Serial Elapsed time", "WdsList": "Wds Fctree Elapsed time"}[method]
        for line in stdout.splitlines():
            if marker in line:
                secs = float(line.split(":")[1].split()[0])
        if secs is None:
            raise RuntimeError(f"reference {method} printed no elapsed time:\n" + stdout[-2000:] + stderr)
        return {"seconds": secs, "peak_rss_mb": rss}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def gpu_time(g, reps: int = 5) -> float:
    """Median ms of hg_decompose on g (graph resident, after a warm-up)."""
    from pargridcpu_amd.hipgrid import Engine

    with Engine.from_graph(g) as eng:
        eng.decompose()
        ms = []
        for _ in range(reps):
            eng.decompose()
            ms.append(eng.timing().decompose_ms)
    return sorted(ms)[len(ms) // 2]


def cpu_baseline(spec, kind: str, kmax_sum: int | None = None, full_entries: int | None = None) -> dict:
    """The reference's ParGrid (built from source into oracle/_ref, -O3) or,
    without it, the oracle's single-thread ParGrid restatement, timed on this
    host:
      * value: on a bounded sample of the workload (the same generator at
        1/100 of n, edges and degree cap, ~10-30 s of CPU work), with the GPU
        timed on the same sample beside it (gpu_same_sample);
      * c2_full: the 1M-vertex, 10-layer workload at full size, reference and
        GPU on the same graph;
      * full_size: the workload itself -- run when it fits the host, else
        reported as not run with the reference's memory estimate.
    cores = host threads the reference ran with (-num_thread = nproc); its
    peel team is fixed at 32 threads whatever -num_thread says
    (FCSyncLeft.cpp:51)."""
    from pargridcpu_amd.graph import C2, MultilayerGraph

    ref_bin = os.path.join(ROOT, "oracle", "_ref", "firmcore_ref")
    if kind == "auto":
        kind = "reference" if os.path.exists(ref_bin) else "port"
    host = host_info()
    threads = int(host["nproc"] or 1)
    scale = min(1.0, 1_000_000 / spec.n)
    sample = dataclasses.replace(spec, n=max(1000, int(spec.n * scale)), m_lo=max(1, int(spec.m_lo * scale)),
                                 m_hi=max(1, int(spec.m_hi * scale)),
                                 cap_deg=spec.cap_deg * scale if spec.cap_deg else 0.0)
    g = MultilayerGraph.generate(sample)
    D = g.directed_entries()
    desc = (f"{'sample of ' + spec.name if scale < 1.0 else spec.name}: n={sample.n} L={sample.L} D={D} (the same "
            f"generator at 1/{round(1 / scale)} of n, edges and degree cap; same gamma, reshuffle and seed)")
    rss = None
    if kind == "reference":
        r = run_reference_pargrid(g, threads)
        secs, rss = r["seconds"], r["peak_rss_mb"]
        desc += (f"; reference ParGrid (main.cpp -m ParGrid, built at -O3; the reference's default build is -O0) "
                 f"with -num_thread {threads}, timed by its own omp_get_wtime bracket")
    else:
        from oracle import oracle as O

        t0 = time.perf_counter()
        O.pargrid(g.n, g.layers())
        secs = time.perf_counter() - t0
        desc += "; oracle ParGrid restatement, single thread"
        threads = 1
    gteps = g.L * D / secs / 1e9
    gms = gpu_time(g)
    out = {"value": round(gteps, 6), "unit": "GTEPS", "cores": threads, "kind": kind, "sample": desc,
           "seconds": round(secs, 4), "peak_rss_mb": rss, "nproc": host["nproc"], "cpu_model": host["cpu_model"],
           "num_thread": threads, "peel_team": 32 if kind == "reference" else 1,
           "gpu_same_sample": {"ms": round(gms, 3), "gteps": round(g.L * D / (gms * 1e-3) / 1e9, 3),
                               "speedup": round(secs * 1e3 / gms, 1)}}
    del g
    if kind == "reference":
        g2 = MultilayerGraph.generate(C2)
        D2 = g2.directed_entries()
        r2 = run_reference_pargrid(g2, threads)
        g2ms = gpu_time(g2)
        out["c2_full"] = {"workload": C2.name, "n": g2.n, "L": g2.L, "directed_entries": D2,
                          "reference_seconds": r2["seconds"], "reference_peak_rss_mb": r2["peak_rss_mb"],
                          "reference_gteps": round(g2.L * D2 / r2["seconds"] / 1e9, 4), "gpu_ms": round(g2ms, 3),
                          "gpu_gteps": round(g2.L * D2 / (g2ms * 1e-3) / 1e9, 3),
                          "speedup": round(r2["seconds"] * 1e3 / g2ms, 1)}
        del g2
    if scale < 1.0:
        # The reference's ParGrid memory on the full workload (FCSyncLeft.cpp):
        # adjacency (4 B per entry + n row pointers and counts per layer,
        # Graph.cpp:17-50), degs (n rows of L ints plus pointers), a deep
        # degs + valid copy per lambda-spine node (serial = 1, :134-152) and
        # an n-byte valid copy per fcgrid cell that is never freed (:10-38).
        n, L = spec.n, spec.L
        cells = kmax_sum if kmax_sum is not None else 0
        ent = full_entries if full_entries is not None else (spec.m_lo + spec.m_hi) * L
        est = (4 * ent + 12 * n * L) + (24 + 4 * L) * n + L * (25 + 4 * L) * n + n * cells
        try:
            mem = os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES")
        except (ValueError, OSError):
            mem = None
        out["full_size"] = {"workload": spec.name, "status": "not run",
                            "reason": (f"reference ParGrid needs about {est / 1e9:.0f} GB of host memory here "
                                       f"({cells} fcgrid cells x {n / 1e6:.0f} MB of never-freed valid copies plus "
                                       f"adjacency, degs and lambda-spine copies); host memory "
                                       f"{(mem or 0) / 1e9:.0f} GB"),
                            "estimated_bytes": int(est), "host_memory_bytes": mem}
    return out


def run_batched(steps: int, count: int = 256, params: str = "tune") -> dict:
    """Batched form: `count` graphs of the batched workload decomposed in one
    persistent launch (hg_batch_*, launch shape from its autotuner, or fixed
    by `params` = "block,unroll,blocks_per_cu,order"), against the same graphs
    decomposed one after another by the single-graph engine (graphs resident,
    one hg_decompose each).  Results must be identical."""
    import numpy as np

    from pargridcpu_amd.graph import MultilayerGraph, batch_workload
    from pargridcpu_amd.hipgrid import Batch, Engine

    t0 = time.perf_counter()
    gs = [MultilayerGraph.generate(s) for s in batch_workload(count)]
    D = sum(g.directed_entries() for g in gs)
    log(f"[bench] batched: generated {count} graphs (D={D}) in {time.perf_counter() - t0:.1f}s")
    b = Batch(gs)
    if params == "tune":
        best, tuned_ms = b.tune(reps=1)
        log(f"[bench] batched: autotuner picked {best} ({tuned_ms:.3f} ms)")
    else:
        block, unroll, bpc, order = (int(x) for x in params.split(","))
        b.set_params(block=block, unroll=unroll, blocks_per_cu=bpc, order=order, use_graph=1)
        best = b.params().as_dict()
    walls = []
    for _ in range(max(steps, 3)):
        w0 = time.perf_counter()
        b.decompose()
        walls.append((time.perf_counter() - w0) * 1e3)
    dev_ms = b.last_ms()[0]
    batch_ms = sorted(walls)[len(walls) // 2]
    engs = [Engine.from_graph(g) for g in gs]
    for e in engs[:8]:
        e.decompose()
    w0 = time.perf_counter()
    for e in engs:
        e.decompose()
    seq_ms = (time.perf_counter() - w0) * 1e3
    exact = all(np.array_equal(b.core_index(i), e.core_index()) for i, e in enumerate(engs))
    for e in engs:
        e.close()
    b.close()
    balg = sum(algorithmic_bytes(g.n, g.L, g.directed_entries()) for g in gs)
    achieved = balg / (dev_ms * 1e-3) / 1e9
    traffic, traffic_src = (measured_traffic("hg_batch_peel_c4", "traffic_bytes_per_launch") if count == 256
                            else (None, None))
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "traffic_source": traffic_src,
            "kernel": "k_batch_pull_heavy + k_batch_pull_light + k_batch_peel (one launch)", "algorithmic_bytes": balg, "avg_ms": round(dev_ms, 4)}
    log(f"[bench] batched: launch {batch_ms:.2f} ms wall ({dev_ms:.2f} ms device), sequential {seq_ms:.1f} ms, "
        f"exact={exact}")
    return {"workload": f"{count} graphs: n log-U[1e3,2e5], L U{{2..16}}, avg deg U[2,20], gamma U[2.1,3.0]",
            "graphs": count, "directed_entries": D, "launch_ms": round(batch_ms, 3), "device_ms": round(dev_ms, 3),
            "graphs_per_s": round(count / (batch_ms * 1e-3), 1), "sequential_ms": round(seq_ms, 3),
            "sequential_graphs_per_s": round(count / (seq_ms * 1e-3), 1), "speedup": round(seq_ms / batch_ms, 2),
            "exact": bool(exact), "params": best, "roofline": roof}


def run_serving(n_requests: int = 160, rates=(20.0, 200.0)) -> dict:
    """Serving form: Poisson arrivals, 80% decompositions and 20%
    densest-subgraph searches; graphs from the batched workload's
    distribution, 10% of them large (n = 5M, L = 12, 60M sampled edges);
    200 GB device-memory budget.  Same arrival trace in every mode, at each
    offered rate (20 req/s: the GPU is not saturated; 200 req/s: arrivals
    outpace one-at-a-time service and the queue builds up):
      continuous          the request loop (4 lanes: requests of at most 4M
                          directed entries batched when 8 or more wait, all
                          others alone on the single-graph engine), every
                          request copied in and uploaded;
      device_queue        8 lanes hand requests of at most 4M entries (up to 8
                          at once) to one device work queue whose running
                          workgroups take their items as they free up; larger
                          ones alone on the single-graph engine;
      one_at_a_time       one request at a time on the single-graph engine;
      resident            the trace's 50 graphs registered first (resident in
                          HBM), requests by handle, the resident
                          decomposition reused by repeat and densest requests;
      resident_recompute  the same, but every request decomposes again;
      resident_device_queue  the same (every request decomposes again), with
                          the small graphs kept as one-graph batches whose
                          items join the device work queue.
    Every answer (core index and densest result) must be identical in all
    modes."""
    import math

    import numpy as np

    from pargridcpu_amd.graph import MultilayerGraph, _unit, batch_spec, large_request_spec
    from pargridcpu_amd.hipgrid import REQ_DECOMPOSE, REQ_DENSEST, Server

    t0 = time.perf_counter()
    small = [MultilayerGraph.generate(batch_spec(r, base=0xC5)) for r in range(48)]
    large = [MultilayerGraph.generate(large_request_spec(r)) for r in range(2)]
    log(f"[bench] serving: generated the request graphs in {time.perf_counter() - t0:.1f}s")

    def make_trace(rate):
        trace, t = [], 0.0
        for r in range(n_requests):
            seed = (0xC5 << 32) | r
            kind = REQ_DENSEST if _unit(seed, 10) < 0.2 else REQ_DECOMPOSE
            g = (large[int(_unit(seed, 12) * len(large))] if _unit(seed, 11) < 0.1
                 else small[int(_unit(seed, 13) * len(small))])
            t += -math.log(1.0 - _unit(seed, 14)) / rate
            trace.append((t, g, kind))
        return trace

    def serve(trace, resident=False, **kw):
        with Server(hbm_budget=200 << 30, **kw) as srv:
            gid = {}
            if resident:  # registered before the clock starts: uploaded once, resident in HBM
                for g in small + large:
                    gid[id(g)] = srv.register(g)
            start = time.perf_counter()
            ids = []
            for ta, g, kind in trace:
                now = time.perf_counter() - start
                if ta > now:
                    time.sleep(ta - now)
                ids.append(srv.submit_resident(gid[id(g)], kind, 1.0) if resident else srv.submit(g, kind, 1.0))
            srv.drain()
            st = srv.stats()
            infos = [srv.wait(i) for i in ids]
            outs = [srv.result(i)[0] for i in ids]
            wds = [(float(i.wds.max_density), i.wds.res_k, i.wds.res_lmd, i.wds.res_len) for i in infos]
        return st, infos, outs, wds

    modes = (("continuous", dict(lanes=4, max_batch=64, large_entries=4_000_000)),
             ("device_queue", dict(lanes=8, max_batch=8, large_entries=4_000_000, device_queue=True)),
             ("one_at_a_time", dict(plain=True)),
             ("resident", dict(lanes=4, resident=True, reuse_results=True)),
             ("resident_recompute", dict(lanes=4, resident=True, reuse_results=False)),
             ("resident_device_queue", dict(lanes=8, max_batch=8, large_entries=4_000_000, resident=True,
                                            reuse_results=False, device_queue=True)))
    by_rate = []
    for rate in rates:
        trace = make_trace(rate)
        res = {"offered_rate": rate}
        outs0 = wds0 = None
        same = True
        for name, kw in modes:
            st, infos, outs, wds = serve(trace, **kw)
            sizes = [i.batch_size for i in infos]
            res[name] = {"p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2),
                         "req_per_s": round(n_requests / st["span_s"], 2), "completed": st["completed"],
                         "failed": st["failed"], "batches": st["batches"],
                         "mean_batch": round(sum(sizes) / len(sizes), 2)}
            log(f"[bench] serving {rate:g} req/s {name}: {res[name]}")
            if outs0 is None:
                outs0, wds0 = outs, wds
            else:
                same = same and all(np.array_equal(a, c) for a, c in zip(outs0, outs)) and wds == wds0
        res["identical"] = same
        c, p = res["continuous"], res["one_at_a_time"]
        res["p99_ratio"] = round(p["p99_ms"] / c["p99_ms"], 2) if c["p99_ms"] else None
        by_rate.append(res)
    return {"requests": n_requests, "by_rate": by_rate,
            "identical": all(r.get("identical", False) for r in by_rate)}


def run_decompose(args, spec) -> dict:
    """The headline form: one full decomposition per step of `spec`, graph
    resident, timed with events on the engine's stream."""
    import torch

    from pargridcpu_amd.graph import MultilayerGraph
    from pargridcpu_amd.hipgrid import Engine

    t0 = time.perf_counter()
    g = MultilayerGraph.generate(spec)
    log(f"[bench] generated {spec.name} in {time.perf_counter() - t0:.1f}s")
    n, L, D = g.n, g.L, g.directed_entries()
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    t0 = time.perf_counter()
    eng = Engine.from_graph(g, stream=stream.cuda_stream)
    upload_s = time.perf_counter() - t0
    log(f"[bench] uploaded n={n} L={L} D={D} in {upload_s:.2f}s ({eng.info()['device_bytes'] / 2**30:.1f} GiB)")
    if args.calib_copy_mib:
        a = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {args.calib_copy_mib} MiB read + {args.calib_copy_mib} MiB written")
        del a, b
    for i in range(args.warmup):
        eng.decompose()
        log(f"[bench] warmup {i}: {eng.timing().decompose_ms:.3f} ms")
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    per_call = []
    start.record(stream)
    w0 = time.perf_counter()
    for i in range(args.steps):
        eng.decompose()
        per_call.append(eng.timing().decompose_ms)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    total_ms = start.elapsed_time(end)
    ms = total_ms / args.steps
    t = eng.timing()
    _, kmax = eng.cells()
    log(f"[bench] steps={args.steps} event total {total_ms:.3f} ms, wall {wall * 1e3:.3f} ms, per-call "
        f"{[round(x, 3) for x in per_call]}, levels={t.levels} launches={t.launches} kmax={kmax.tolist()}")
    gteps = L * D / (ms * 1e-3) / 1e9
    balg = algorithmic_bytes(n, L, D)
    dev_ms = sum(per_call) / len(per_call)
    achieved = balg / (dev_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(f"hg_decompose_{args.workload}")
    l2_hits, _ = measured_traffic(f"hg_decompose_{args.workload}", "l2_hits_per_decomposition")
    l2_miss, _ = measured_traffic(f"hg_decompose_{args.workload}", "l2_misses_per_decomposition")
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "traffic_source": traffic_src,
            "counted_hbm_gbs": round(traffic / (dev_ms * 1e-3) / 1e9, 1) if traffic else None,
            "l2_hit_rate": round(l2_hits / (l2_hits + l2_miss), 4) if l2_hits is not None and l2_miss else None,
            "kernel": "hg_decompose (k_init + per-level k_scan/k_cascade + binned dense rounds k_bin_expand/scatter/group/fine)", "algorithmic_bytes": balg,
            "avg_ms": round(dev_ms, 4)}
    eng.close()
    del eng, g
    cpu = None
    if args.cpu_baseline != "off":
        try:
            cpu = cpu_baseline(spec, args.cpu_baseline, kmax_sum=int(kmax.sum()), full_entries=D)
            log(f"[bench] cpu baseline: {cpu}")
        except Exception as e:  # reported, never fatal for the GPU number
            cpu = {"value": None, "unit": "GTEPS", "cores": None, "kind": args.cpu_baseline, "sample": f"failed: {e}"}
    return {"value": round(gteps, 4), "ms_per_step": round(ms, 4), "roofline": roof, "cpu_baseline": cpu,
            "config": {"workload": spec.name, "n": n, "L": L, "directed_entries": D, "gamma": spec.gamma,
                       "rho": spec.rho, "offsets": "int64", "parallelism": "single-gpu"}}


def host_table() -> dict:
    """Rows of the CPU-baseline table on this host for the 20k- and
    1M-vertex workloads at full size: the reference's ParGrid (nproc
    threads, peel team 32), its serial lpe and kpe, WdsList at beta = 1,
    each with its own elapsed-time line and peak RSS; the GPU engine on the
    same graphs (decomposition and densest subgraph)."""
    from pargridcpu_amd.graph import C1, C2, MultilayerGraph
    from pargridcpu_amd.hipgrid import Engine

    host = host_info()
    threads = int(host["nproc"] or 1)
    rows = []
    for spec in (C1, C2):
        g = MultilayerGraph.generate(spec)
        D = g.directed_entries()
        for method, extra in (("ParGrid", ()), ("lpe", ()), ("kpe", ()), ("WdsList", ("-beta", "1.0"))):
            r = run_reference_pargrid(g, threads, method, extra)
            rows.append({"workload": spec.name, "method": f"reference {method}", "threads": threads,
                         "seconds": r["seconds"], "peak_rss_mb": r["peak_rss_mb"],
                         "gteps": round(g.L * D / r["seconds"] / 1e9, 4) if method != "WdsList" else None})
            log(f"[bench] host table: {rows[-1]}")
        with Engine.from_graph(g) as eng:
            eng.decompose()
            ms, dms = [], []
            for _ in range(5):
                eng.decompose()
                ms.append(eng.timing().decompose_ms)
                eng.densest(1.0)
                dms.append(eng.timing().densest_ms)
            dev = eng.info()["device_bytes"]
        t = sorted(ms)[2]
        rows.append({"workload": spec.name, "method": "HipGrid hg_decompose", "seconds": round(t / 1e3, 6),
                     "gteps": round(g.L * D / (t * 1e-3) / 1e9, 3), "device_bytes": dev})
        rows.append({"workload": spec.name, "method": "HipGrid hg_densest (beta 1.0)",
                     "seconds": round(sorted(dms)[2] / 1e3, 6)})
        log(f"[bench] host table: {rows[-2]} {rows[-1]}")
    return {"host": host, "rows": rows}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=os.environ.get("HG_WORKLOAD", "c3"))
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "reference", "port", "off"])
    ap.add_argument("--decompose", type=int, default=1, help="run the headline decomposition (0: skip, profiling)")
    ap.add_argument("--batched", type=int, default=1, help="also run the batched form (0: skip)")
    ap.add_argument("--batch-params", default="tune",
                    help='batched launch shape: "tune" or "block,unroll,blocks_per_cu,order"')
    ap.add_argument("--serving", type=int, default=1, help="also run the serving form (0: skip)")
    ap.add_argument("--host-table", action="store_true",
                    help="only print the CPU-baseline table of the 20k- and 1M-vertex workloads (reference modes and "
                         "the engine, full size) as JSON")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the lambda-sharded multi-GPU "
              f"layout is a design, not built", file=sys.stderr)
        return 2

    if args.host_table:
        print(json.dumps(host_table()), flush=True)
        return 0

    from pargridcpu_amd.graph import WORKLOADS

    spec = WORKLOADS[args.workload]
    head = {"value": None, "ms_per_step": None, "roofline": None, "cpu_baseline": None,
            "config": {"workload": spec.name}}
    if args.decompose:
        head = run_decompose(args, spec)
    batched = serving = None
    if args.batched:
        try:
            batched = run_batched(args.steps, params=args.batch_params)
        except Exception as e:  # reported, never fatal for the headline number
            batched = {"error": str(e)}
    if args.serving:
        try:
            serving = run_serving()
        except Exception as e:
            serving = {"error": str(e)}
    line = {
        "metric": METRIC, "value": head["value"], "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded Chung-Lu multilayer generator)",
        "config": head["config"], "roofline": head["roofline"], "cpu_baseline": head["cpu_baseline"],
        "batched": batched, "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
