"""Benchmark of the MI355X simplifier, next to the reference CPU library on the same host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload patches4096|icosphere8|grid256|seamed1024|serving|sloppy|update|clod]

Default (patches4096): the batched form.  A step is one meshopt_hip_simplify_batch_device call that
simplifies 4096 clusterlod-style patches (64-512 triangles each, LockBorder, target 50%, error 1e-2)
with every array resident in HBM; the line reports input Mtri/s, bit-exactness of every patch against
the reference library, the reference library on one host core and on all of the host's cores (a
thread pool, one patch per work item), the plain per-call device path on the same patches, and a
nested "plain" object: the plain meshopt_simplify call on the icosphere L8 mesh (device-resident and
through the host entry point, H2D/D2H included).  The plain workloads alone: --workload icosphere8
and the others.  The timed region is bracketed by device synchronisation.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np

HBM_PEAK_GBS = 8000.0



def make_workload(name):
    from meshoptimizer_amd import workloads as W
    if name == "icosphere8":
        m = W.icosphere(8)
        return m, W.target_count(m.index_count, 1, 10), None, "icosphere L8 (655,362 V / 1,310,720 T), target 10%, error 1e-2, options 0"
    if name == "grid256":
        m = W.grid(256)
        return m, W.target_count(m.index_count, 1, 4), None, "grid 256x256 (65,536 V / 130,050 T), target 25%, error 1e-2, options 0"
    if name == "seamed1024":
        m = W.seamed_grid(1024)
        return m, W.target_count(m.index_count, 1, 5), W.SEAMED_GRID_WEIGHTS, "grid 1024x1024 + normal/UV, 16 seams, target 20%, error 1e-2"
    raise SystemExit(f"unknown workload {name}")


def alg_bytes(T0, V, A, passes):
    """Compulsory HBM traffic of one call (fixed formula): setup plus, per pass,
    96 T_p + 12 T_{p+1} + 64 C_p + (109 + W) V + (88 + [A>0](88 + 32A)) M_p."""
    has_a = 1 if A else 0
    W_ = 56 + has_a * (88 + 20 * A)
    total = (120 + 12 * has_a) * T0 + (338 + has_a * (56 + 28 * A)) * V
    for i, p in enumerate(passes):
        Tp = p["live_triangles"]
        Tn = passes[i + 1]["live_triangles"] if i + 1 < len(passes) else Tp
        total += 96 * Tp + 12 * Tn + 64 * p["candidates"] + (109 + W_) * V + (88 + has_a * (88 + 32 * A)) * p["committed"]
    return total


def perform_bytes(V, passes):
    """The perform term of the fixed formula (the commit of the sorted candidates, which the
    commit rounds carry): 16 C_p + 24 T_p + 35 V per pass."""
    return sum(16 * p["candidates"] + 24 * p["live_triangles"] + 35 * V for p in passes)


def batch_alg_bytes(index_counts, vertex_counts, out_counts, stats):
    """The fixed formula summed over the meshes of a batch (no attributes, so W = 56): setup
    120 T_0 + 338 V, and per pass 96 T_p + 12 T_{p+1} + 64 C_p + 165 V + 88 M_p.  stats[p] holds the
    mesh's passes and its sums of T_p, C_p and M_p; the sum of T_{p+1} is sum T_p - T_0 + T_final."""
    T0 = index_counts.astype(np.int64) // 3
    V = vertex_counts.astype(np.int64)
    Tf = out_counts.astype(np.int64) // 3
    st = stats.astype(np.int64)
    passes, sT, sC, sM = st[:, 0], st[:, 1], st[:, 2], st[:, 3]
    sTn = np.where(passes > 0, sT - T0 + Tf, 0)
    total = 120 * T0 + 338 * V + 96 * sT + 12 * sTn + 64 * sC + 165 * V * passes + 88 * sM
    return int(total.sum())


def host_description():
    """CPU model, logical CPUs, and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def measured_traffic(workload, kernel="k_round_x"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary of this workload
    (FETCH_SIZE and WRITE_SIZE runs, calibrated against a copy of known size), or None."""
    import re
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"Traffic per launch = ([0-9.]+) MB", open(path).read())
    return float(m.group(1)) * 1e6 if m else None


def cpu_reference(m, target, weights, budget_s=12.0):
    """The reference library (compiled from its own sources into oracle/_ref) on one host core."""
    from oracle import oracle as O
    kind = "reference" if O.have_ref() else "port"
    impl = "ref" if O.have_ref() else "oracle"
    A = m.vertices[:, 3:3 + len(weights)] if weights is not None else None
    times = []
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        O.simplify(m.indices, m.vertices, target, 1e-2, 0, attributes=A, weights=weights, impl=impl)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > budget_s or len(times) >= 10:
            break
    med = statistics.median(times)
    return {"value": m.triangle_count / med / 1e6, "unit": "Mtri/s", "cores": 1, "kind": kind,
            "sample": f"{len(times)} single-threaded calls on the same mesh, median {med * 1e3:.1f} ms"}


def calibrate_copy(torch, dev):
    """One device copy of 256 MiB: the reference point for the PMC byte counters (its FETCH_SIZE
    and WRITE_SIZE should read 262144 KiB)."""
    src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def batch_compulsory_bytes(index_counts, vertex_counts, out_counts):
    """HBM bytes a batched call cannot avoid: indices read (4 I) and positions read (12 V, xyz), the
    result written (4 I_out), offsets/target read and count/error/status written (6 x 4 per mesh)."""
    return int(4 * np.sum(index_counts) + 12 * np.sum(vertex_counts) + 4 * np.sum(out_counts) + 24 * len(index_counts))


def bench_patches(args):
    """The batched workload: 4096 clusterlod-style patches in one meshopt_hip_simplify_batch_device
    call per step."""
    import torch
    import meshoptimizer_amd as M
    from meshoptimizer_amd import workloads as W

    count, opts, err = 4096, M.SimplifyLockBorder, 1e-2
    meshes = W.patches(count=count)
    idx, io, verts, vo, stride = M.pack_meshes([(m.indices, m.vertices) for m in meshes])
    targets = np.array([W.target_count(m.index_count, 1, 2) for m in meshes], np.uint32)
    tris = int(sum(m.triangle_count for m in meshes))
    desc = (f"{count} clusterlod-style patches of 64-512 triangles ({tris:,} T) cut from a displaced 2048x2048 grid, "
            "LockBorder, target 50% per patch, error 1e-2, one batched call")

    dev = torch.device("cuda:0")
    as_dev = lambda a, dt: torch.from_numpy(a.view(dt)).to(dev)
    d_idx = as_dev(idx, np.int32)
    d_io = as_dev(io, np.int32)
    d_vo = as_dev(vo, np.int32)
    d_tg = as_dev(targets, np.int32)
    d_v = torch.from_numpy(verts).to(dev)
    d_dst = torch.empty(max(idx.size, 1), dtype=torch.int32, device=dev)
    d_cnt = torch.empty(count, dtype=torch.int32, device=dev)
    d_err = torch.empty(count, dtype=torch.float32, device=dev)

    def step():
        M.simplify_batch_device(d_dst.data_ptr(), d_cnt.data_ptr(), d_err.data_ptr(), d_idx.data_ptr(), d_io.data_ptr(),
                                d_v.data_ptr(), d_vo.data_ptr(), int(vo[-1]), stride, d_tg.data_ptr(), count, err, opts)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if args.calibrate_copy:
        calibrate_copy(torch, dev)
    step()
    torch.cuda.synchronize()
    g_dst = d_dst.cpu().numpy().view(np.uint32)
    g_cnt = d_cnt.cpu().numpy().view(np.uint32)
    g_err = d_err.cpu().numpy()

    times = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        times.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    value = tris / (elapsed / args.steps) / 1e6

    # the batched call's own span from HIP events on the library's stream, in separate calls
    # (the events add host work to every call, so they stay out of the timed ones)
    M.kernel_timing(True)
    M.batch_kernel_timing()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    k_ms, k_launches, k_meshes = M.batch_kernel_timing()
    M.kernel_timing(False)

    # the plain per-call device path on the same patches (a bounded sample, per-patch mean scaled)
    sample = min(count, args.plain_sample)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p in range(sample):
        M.simplify_device(d_dst.data_ptr() + 4 * int(io[p]), d_idx.data_ptr() + 4 * int(io[p]), int(io[p + 1] - io[p]),
                          d_v.data_ptr() + stride * int(vo[p]), int(vo[p + 1] - vo[p]), stride, int(targets[p]), err, opts)
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t0) / sample * count
    sample_tris = int(sum(meshes[p].triangle_count for p in range(sample)))
    plain_value = sample_tris / (plain_s * sample / count) / 1e6

    stats = M.last_batch_stats(count)

    # the launch shape in use (a tuned table, or the default) and, if tuned, the default's time
    shape = M.batch_get_shape()
    default_shape = {"per_cu": [6, 5, 2, 1], "threads": [256, 256, 256, 1024], "cursor": [0, 0, 0, 1]}
    tuned = any(shape[k] != v for k, v in default_shape.items())
    default_ms = None
    if tuned:
        M.batch_set_shape()
        for _ in range(2):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        default_ms = (time.perf_counter() - t0) / args.steps * 1e3
        M.batch_set_shape(shape["per_cu"], shape["threads"], shape["cursor"])
    out = {
        "metric": "input Mtri/s simplified (bit-exact vs CPU), batched vs plain; p99 latency",
        "value": round(value, 3),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": desc, "input_triangles": tris, "output_triangles": int(g_cnt.sum()) // 3, "patches": count,
                   "passes": int(stats[:, 0].sum()), "candidates": int(stats[:, 2].sum()), "committed": int(stats[:, 3].sum())},
        "p50_ms": round(statistics.median(times) * 1e3, 4),
        "p99_ms": round(float(np.percentile(np.array(times) * 1e3, 99)), 4),
        "plain_per_call_gpu": {"value": round(plain_value, 3), "unit": "Mtri/s",
                               "sample": f"meshopt_hip_simplify_device on the first {sample} patches, one call each",
                               "batch_speedup": round(value / plain_value, 2)},
        "batch_shape": {**{k: shape[k] for k in ("per_cu", "threads", "cursor", "index_caps", "vertex_caps")},
                        "source": "tuned table (meshoptimizer_amd/tune.py)" if tuned else "default",
                        "default_shape_ms_per_step": round(default_ms, 4) if default_ms else None,
                        "gain_over_default": round(default_ms / ms, 4) if default_ms else None},
    }
    if k_launches:
        ab = batch_alg_bytes(np.diff(io), np.diff(vo), g_cnt, stats)
        per_launch_s = k_ms * 1e-3 / k_launches
        achieved = ab / per_launch_s / 1e9
        comp = batch_compulsory_bytes(np.diff(io), np.diff(vo), g_cnt)
        out["roofline"] = {"bound": "hbm",
                           "kernel": "one batched call: k_batch_classify_order (classification and largest-first order) and the "
                                     "k_simplify_patches size classes launched side by side, timed with HIP events from the "
                                     "classification start to the join of the class streams",
                           "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 6),
                           "traffic": measured_traffic("patches4096", "k_simplify_patches"),
                           "alg_bytes_per_launch": ab, "avg_launch_us": round(per_launch_s * 1e6, 2),
                           "alg_bytes_formula": "per patch: 120 T0 + 338 V + sum over passes (96 T_p + 12 T_p+1 + 64 C_p + 165 V + 88 M_p)",
                           "compulsory_hbm_bytes": comp,
                           "compulsory_hbm_frac": round(comp / per_launch_s / 1e9 / HBM_PEAK_GBS, 6)}

    if not args.no_cpu:
        from oracle import oracle as O
        impl = "ref" if O.have_ref() else "oracle"
        kind = "reference" if impl == "ref" else "port"
        r_dst, r_cnt, r_err = O.simplify_batch(idx, io, verts, vo, targets, err, opts, threads=1, impl=impl)
        exact = bool(np.array_equal(r_cnt, g_cnt) and np.array_equal(r_err.view(np.uint32), g_err.view(np.uint32)))
        for p in range(count):
            if not exact:
                break
            exact = np.array_equal(r_dst[io[p]:io[p] + r_cnt[p]], g_dst[io[p]:io[p] + g_cnt[p]])
        out["bit_exact"] = bool(exact)
        host = host_description()
        cores = max(1, min(args.cpu_threads, host["affinity_cpus"]))
        base = {}
        for th in (1, cores):
            reps, t_all = 0, 0.0
            while t_all < args.cpu_budget / 2 or reps < 1:
                t0 = time.perf_counter()
                O.simplify_batch(idx, io, verts, vo, targets, err, opts, threads=th, impl=impl)
                t_all += time.perf_counter() - t0
                reps += 1
            base[th] = (tris * reps / t_all / 1e6, reps, t_all)
        v1, vN = base[1], base[cores]
        out["cpu_baseline"] = {"value": round(vN[0], 3), "unit": "Mtri/s", "cores": cores, "kind": kind,
                               "sample": f"the full 4096-patch batch, {vN[1]} runs on {cores} threads (std::thread pool, one patch per "
                                         f"work item from an atomic index; {cores} = the CPUs this process may use, at most the GPU's "
                                         f"16-CPU share of the host)",
                               "single_thread": round(v1[0], 3), "single_thread_sample": f"{v1[1]} runs of the full batch on 1 thread",
                               "host": host}
        out["speedup_vs_cpu_all_cores"] = round(value / vN[0], 2)
        out["speedup_vs_cpu_one_core"] = round(value / v1[0], 2)
    return out


def replay_cpu(reqs, arrivals, threads):
    """The reference library (oracle/_ref, or the restatement) serving the same arrival trace on a
    pool of `threads` workers; returns ([(indices, error)], [latency s], makespan s)."""
    import concurrent.futures as cf
    import threading
    from oracle import oracle as O
    impl = "ref" if O.have_ref() else "oracle"
    out = [None] * len(reqs)
    lat = [0.0] * len(reqs)
    t0 = time.perf_counter()

    def run(k):
        r = reqs[k]
        m = r.mesh
        attrs = m.vertices[:, 3:3 + m.attr_count] if r.weights is not None else None
        out[k] = O.simplify(m.indices, m.vertices, r.target_index_count, 1e-2, 0, attributes=attrs, weights=r.weights, impl=impl)
        lat[k] = time.perf_counter() - t0 - arrivals[k]

    futs = []
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        for k in range(len(reqs)):
            delay = arrivals[k] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            futs.append(ex.submit(run, k))
        for f in futs:
            f.result()
    return out, lat, time.perf_counter() - t0 - arrivals[0]


def bench_serving(args):
    """The serving stream: Poisson arrivals of requests of 1k-2M triangles (`--min-tris`, `--max-tris`; grids and icospheres,
    a third with normal+UV attributes, target ratio 0.5/0.25/0.1) served by meshopt_hip_server
    with `--slots` concurrent device contexts and a 64 GB admission budget.  First the closed-loop
    capacity (every request queued at once), then an open-loop replay at `--load` of it."""
    import threading
    import torch
    import meshoptimizer_amd as M
    from meshoptimizer_amd import workloads as W

    count, slots, budget, load = args.requests, args.slots, 64 << 30, args.load
    reqs = W.request_stream(count, 1.0, tmin=args.min_tris, tmax=args.max_tris)
    unit = np.array([r.arrival_s for r in reqs])  # arrival times at one request per second
    tris = np.array([r.mesh.triangle_count for r in reqs], dtype=np.float64)
    total = float(tris.sum())
    torch.cuda.init()

    srv = M.Server(budget_bytes=budget, slots=slots)

    def serve(arrivals):
        res = [None] * count
        tickets = [None] * count
        done = threading.Event()

        def waiter():
            for k in range(count):
                while tickets[k] is None:
                    time.sleep(0.0002)
                res[k] = srv.wait(tickets[k])
            done.set()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for k, r in enumerate(reqs):
            delay = arrivals[k] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            m = r.mesh
            attrs = m.vertices[:, 3:3 + m.attr_count] if r.weights is not None else None
            tickets[k] = srv.submit(m.indices, m.vertices, r.target_index_count, 1e-2, 0, attributes=attrs, weights=r.weights)
        done.wait()
        th.join()
        return res

    serve(np.zeros(count))  # warm-up: every slot's arena grows to its working size
    closed = serve(np.zeros(count))
    cap_s = (max(i["finish_ms"] for _, _, i in closed) - min(i["submit_ms"] for _, _, i in closed)) * 1e-3
    capacity = total / cap_s
    rate = load * capacity / tris.mean()  # requests per second
    arrivals = unit / rate
    opened = serve(arrivals)
    srv.close()
    lat_ms = np.array([i["finish_ms"] - i["submit_ms"] for _, _, i in opened])
    span_s = (max(i["finish_ms"] for _, _, i in opened) - min(i["submit_ms"] for _, _, i in opened)) * 1e-3
    value = total / span_s / 1e6

    out = {
        "metric": "input Mtri/s served (bit-exact vs CPU), serving stream",
        "value": round(value, 3),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(span_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"serving stream: {count} requests, log-uniform {args.min_tris:g}-{args.max_tris:g} triangles (grids and icospheres, "
                               f"1/3 with normal+UV), target ratio 0.5/0.25/0.1, Poisson arrivals at {load:.0%} of the "
                               f"measured capacity, {slots} slots, 64 GB admission budget",
                   "requests": count, "input_triangles": int(total), "slots": slots,
                   "offered_requests_per_s": round(rate, 2),
                   "requests_batched": int(sum(1 for _, _, i in opened if i["batched"] > 1))},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "closed_loop_capacity": round(capacity / 1e6, 3),
    }
    if not args.no_cpu:
        cores = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        cpu_res, cpu_lat, cpu_span = replay_cpu(reqs, arrivals, cores)
        exact = all(np.array_equal(g[0], c[0]) and np.float32(g[1]).tobytes() == np.float32(c[1]).tobytes()
                    for g, c in zip(opened, cpu_res))
        out["bit_exact"] = bool(exact)
        from oracle import oracle as O
        out["cpu_baseline"] = {"value": round(total / cpu_span / 1e6, 3), "unit": "Mtri/s", "cores": cores,
                               "kind": "reference" if O.have_ref() else "port",
                               "sample": f"the same {count}-request arrival trace on {cores} worker threads",
                               "p50_ms": round(float(np.percentile(np.array(cpu_lat) * 1e3, 50)), 3),
                               "p99_ms": round(float(np.percentile(np.array(cpu_lat) * 1e3, 99)), 3)}
    print(json.dumps(out), flush=True)


def bench_sloppy(args):
    """meshopt_simplifySloppy on the icosphere L8 mesh (target 10%, error 1e-2) with inputs resident in
    HBM: one step is one whole call (the grid-size search, cells, cell quadrics, filtering)."""
    import torch
    import meshoptimizer_amd as M
    from meshoptimizer_amd import workloads as W

    m = W.icosphere(8)
    target = W.target_count(m.index_count, 1, 10)
    dev = torch.device("cuda:0")
    ib = torch.from_numpy(m.indices.astype(np.int32)).to(dev)
    vb = torch.from_numpy(np.ascontiguousarray(m.vertices[:, :3])).to(dev)
    dest = torch.empty(m.index_count, dtype=torch.int32, device=dev)

    def step():
        return M.simplify_sloppy_device(dest.data_ptr(), ib.data_ptr(), m.index_count, vb.data_ptr(), m.vertex_count, 12, None,
                                        target, 1e-2)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    n_gpu, e_gpu = step()
    gpu_out = dest[:n_gpu].cpu().numpy().astype(np.uint32)
    times = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        times.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    out = {
        "metric": "input Mtri/s simplified (bit-exact vs CPU), meshopt_simplifySloppy",
        "value": round(m.triangle_count / (elapsed / args.steps) / 1e6, 3),
        "unit": "Mtri/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded generator)",
        "config": {"workload": "meshopt_simplifySloppy, icosphere L8 (655,362 V / 1,310,720 T), target 10%, error 1e-2",
                   "input_triangles": m.triangle_count, "output_triangles": n_gpu // 3},
        "p50_ms": round(statistics.median(times) * 1e3, 4),
        "p99_ms": round(float(np.percentile(np.array(times) * 1e3, 99)), 4),
    }
    if not args.no_cpu:
        from oracle import oracle as O
        impl = "ref" if O.have_ref() else "oracle"
        v3 = np.ascontiguousarray(m.vertices[:, :3])
        ro, eo = O.simplify_sloppy(m.indices, v3, target, 1e-2, impl=impl)
        out["bit_exact"] = bool(np.array_equal(ro, gpu_out) and np.float32(eo).tobytes() == np.float32(e_gpu).tobytes())
        cpu_t = []
        t_start = time.perf_counter()
        while time.perf_counter() - t_start < min(args.cpu_budget, 5.0) and len(cpu_t) < 20:
            ts = time.perf_counter()
            O.simplify_sloppy(m.indices, v3, target, 1e-2, impl=impl)
            cpu_t.append(time.perf_counter() - ts)
        med = statistics.median(cpu_t)
        out["cpu_baseline"] = {"value": round(m.triangle_count / med / 1e6, 3), "unit": "Mtri/s", "cores": 1,
                               "kind": "reference" if impl == "ref" else "port",
                               "sample": f"{len(cpu_t)} single-threaded calls on the same mesh, median {med * 1e3:.2f} ms"}
    print(json.dumps(out), flush=True)


def bench_update(args):
    """meshopt_simplifyWithUpdate on the seamed 1024x1024 grid with its normal/UV attributes (target
    20%, error 1e-2), inputs resident in HBM: one step is one whole call, the vertex solve and
    write-back included.  The call updates its arrays in place, so every step starts from a pristine
    device copy restored (and synchronised) outside the timed call."""
    import torch
    import meshoptimizer_amd as M

    m, target, weights, desc = make_workload("seamed1024")
    dev = torch.device("cuda:0")
    ib0 = torch.from_numpy(m.indices.astype(np.int32)).to(dev)
    vb0 = torch.from_numpy(m.vertices).to(dev)
    ib, vb = ib0.clone(), vb0.clone()
    dest = torch.empty(m.index_count, dtype=torch.int32, device=dev)
    stride = m.stride

    def restore():
        ib.copy_(ib0)
        vb.copy_(vb0)
        torch.cuda.synchronize()

    def step():
        return M.simplify_with_update_device(ib.data_ptr(), m.index_count, vb.data_ptr(), m.vertex_count, stride, target, 1e-2, 0,
                                             attributes_ptr=vb.data_ptr() + 12, attribute_stride=stride, weights=weights)

    def collapse_only():
        return M.simplify_device(dest.data_ptr(), ib0.data_ptr(), m.index_count, vb0.data_ptr(), m.vertex_count, stride, target, 1e-2, 0,
                                 attributes_ptr=vb0.data_ptr() + 12, attribute_stride=stride, weights=weights)

    for _ in range(args.warmup):
        restore()
        step()
    restore()
    n_gpu, e_gpu = step()
    gpu_ib = ib[:n_gpu].cpu().numpy().astype(np.uint32)
    gpu_vb = vb.cpu().numpy()
    call, _ = M.last_call_stats()

    times, plain = [], []
    for _ in range(args.steps):
        restore()
        ts = time.perf_counter()
        step()
        times.append(time.perf_counter() - ts)
    for _ in range(args.steps):
        ts = time.perf_counter()
        collapse_only()
        plain.append(time.perf_counter() - ts)
    mean = sum(times) / len(times)
    out = {
        "metric": "input Mtri/s simplified and solved (bit-exact vs CPU), meshopt_simplifyWithUpdate",
        "value": round(m.triangle_count / mean / 1e6, 3),
        "unit": "Mtri/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(mean * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded generator)",
        "config": {"workload": "meshopt_simplifyWithUpdate, " + desc, "input_triangles": m.triangle_count,
                   "output_triangles": n_gpu // 3, "vertices_left_to_rounds": call["solve_vertices"],
                   "solve_rounds_launched": call["solve_rounds"]},
        "p50_ms": round(statistics.median(times) * 1e3, 4),
        "p99_ms": round(float(np.percentile(np.array(times) * 1e3, 99)), 4),
        "same_call_without_solve_ms": round(statistics.median(plain) * 1e3, 4),
    }
    if not args.no_cpu:
        from oracle import oracle as O
        impl = "ref" if O.have_ref() else "oracle"
        na = len(weights)

        def cpu_call():
            b = m.vertices.copy()
            r, e = O.simplify_with_update(m.indices, b, target, 1e-2, 0, attributes=b[:, 3:3 + na], weights=weights, impl=impl)
            return r, e, b

        ts = time.perf_counter()
        ro, eo, bo = cpu_call()
        cpu_t = [time.perf_counter() - ts]
        out["bit_exact"] = bool(np.array_equal(ro, gpu_ib) and np.float32(eo).tobytes() == np.float32(e_gpu).tobytes()
                                and bo.tobytes() == gpu_vb.tobytes())
        t_start = time.perf_counter()
        while time.perf_counter() - t_start < min(args.cpu_budget, 10.0) and len(cpu_t) < 5:
            ts = time.perf_counter()
            cpu_call()
            cpu_t.append(time.perf_counter() - ts)
        med = statistics.median(cpu_t)
        out["cpu_baseline"] = {"value": round(m.triangle_count / med / 1e6, 3), "unit": "Mtri/s", "cores": 1,
                               "kind": "reference" if impl == "ref" else "port",
                               "sample": f"{len(cpu_t)} single-threaded calls on the same mesh, median {med * 1e3:.1f} ms"}
    print(json.dumps(out), flush=True)


def clod_level_workload(n=1024, tiles=64):
    """One clusterlod level: a displaced n x n grid's triangles split into tiles x tiles spatial
    groups (about 512 triangles each at the defaults), the position remap, no caller locks."""
    from meshoptimizer_amd import workloads as W
    m = W.grid(n)
    v = np.ascontiguousarray(m.vertices[:, :3])
    tri = m.indices.reshape(-1, 3)
    c = v[tri].mean(axis=1)
    g = np.minimum((c[:, 1] * tiles).astype(np.int64), tiles - 1) * tiles + np.minimum((c[:, 0] * tiles).astype(np.int64), tiles - 1)
    order = np.argsort(g, kind="stable")
    counts = np.bincount(g, minlength=tiles * tiles)
    io = np.zeros(tiles * tiles + 1, np.uint32)
    io[1:] = np.cumsum(counts) * 3
    idx = np.ascontiguousarray(tri[order].reshape(-1)).astype(np.uint32)
    remap = np.arange(v.shape[0], dtype=np.uint32)  # a grid has no repeated positions
    targets = (np.diff(io) // 3 // 2 * 3).astype(np.uint32)
    return m, v, idx, io, remap, targets


def bench_clod(args):
    """One clusterlod level on the device (meshopt_hip_clod_level_device): boundary locks, the
    packing of 4096 groups of about 512 triangles cut from a 1024 x 1024 grid, their batched
    simplification to half (ErrorAbsolute | Permissive, no error limit, with clodDefaultConfig's
    sloppy fallback, as clodBuild calls it) and the mapping back; inputs resident in HBM.  The
    reference runs the same level as clodBuild does: its lockBoundary, then its simplify() per
    group (one meshopt_simplifyWithAttributes(Sparse), and the fallback for a group left above its
    target)."""
    import torch
    import meshoptimizer_amd as M

    m, v, idx, io, remap, targets = clod_level_workload(tiles=args.clod_tiles)
    G = targets.size
    flt_max = float(np.finfo(np.float32).max)
    opts = M.SimplifyErrorAbsolute | M.SimplifyPermissive
    # clodDefaultConfig (demo/clusterlod.h:573-580): Permissive in the options, sloppy fallback
    config = {"fallback_permissive": 0, "fallback_sloppy": 1, "error_factor_sloppy": 2.0, "error_edge_limit": 0.0}
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).to(dev)
    d_idx, d_io, d_remap, d_tg, d_v = t(idx), t(io), t(remap), t(targets), torch.from_numpy(v).to(dev)
    d_locks = torch.empty(v.shape[0], dtype=torch.uint8, device=dev)
    d_dst = torch.empty(idx.size, dtype=torch.int32, device=dev)
    d_cnt = torch.empty(G, dtype=torch.int32, device=dev)
    d_err = torch.empty(G, dtype=torch.float32, device=dev)

    def step():
        M.clod_level_device(d_dst.data_ptr(), d_cnt.data_ptr(), d_err.data_ptr(), d_idx.data_ptr(), d_io.data_ptr(), G, d_v.data_ptr(),
                            v.shape[0], 12, d_remap.data_ptr(), None, d_locks.data_ptr(), d_tg.data_ptr(), flt_max, opts, config)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        times.append(time.perf_counter() - ts)
    mean = sum(times) / len(times)
    out = {
        "metric": "input Mtri/s simplified (bit-exact vs CPU), one clusterlod level on the device",
        "value": round(m.triangle_count / mean / 1e6, 3),
        "unit": "Mtri/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(mean * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded generator)",
        "config": {"workload": f"clusterlod level: grid 1024x1024 ({m.triangle_count:,} T) in {G} spatial groups "
                               f"(about {m.triangle_count // G} T each), boundary locks, each group to 50%, ErrorAbsolute | Permissive, "
                               "no error limit, sloppy fallback (clodDefaultConfig)", "groups": int(G)},
        "p50_ms": round(statistics.median(times) * 1e3, 4),
        "p99_ms": round(float(np.percentile(np.array(times) * 1e3, 99)), 4),
    }
    if not args.no_cpu:
        from oracle import oracle as O
        from oracle import clod as C
        impl = "ref" if O.have_ref() else "oracle"
        counts = d_cnt.cpu().numpy().astype(np.uint32)
        dst = d_dst.cpu().numpy().astype(np.uint32)
        errs = d_err.cpu().numpy()
        uses = np.zeros(v.shape[0], np.int32)
        for p in range(G):
            uses[np.unique(remap[idx[io[p]:io[p + 1]]])] += 1
        locks = (uses >= 2)[remap].astype(np.uint8)
        exact = bool(np.array_equal(d_locks.cpu().numpy(), locks))
        # a bounded, evenly spaced sample of groups, timed and checked
        sample = list(range(0, G, max(1, G // 256)))
        cpu_t, tris = 0.0, 0
        for p in sample:
            grp = idx[io[p]:io[p + 1]]
            ts = time.perf_counter()
            ro, eo = C.clod_simplify(grp, v, locks, int(targets[p]), opts, config, impl=impl)
            cpu_t += time.perf_counter() - ts
            tris += grp.size // 3
            exact &= bool(np.array_equal(dst[io[p]:io[p] + counts[p]], ro) and np.float32(eo).tobytes() == np.float32(errs[p]).tobytes())
        out["bit_exact"] = exact
        out["cpu_baseline"] = {"value": round(tris / cpu_t / 1e6, 3), "unit": "Mtri/s", "cores": 1,
                               "kind": "reference" if impl == "ref" else "port",
                               "sample": f"{len(sample)} of {G} groups, clodBuild's simplify() each on one thread ({cpu_t * 1e3:.0f} ms); "
                                         "every sampled group checked bit-exact"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="patches4096")
    ap.add_argument("--no-plain", action="store_true", help="patches4096: leave out the nested plain icosphere L8 line")
    ap.add_argument("--clod-tiles", type=int, default=64,
                    help="clod workload: the grid is cut into tiles x tiles groups (64: about 512 triangles each; 34: about "
                         "1810, sixteen clusters of the clusterlod demo's default size)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=256, help="patches run through the plain per-call path")
    ap.add_argument("--requests", type=int, default=64, help="serving: requests in the stream")
    ap.add_argument("--slots", type=int, default=4, help="serving: concurrent device contexts")
    ap.add_argument("--load", type=float, default=0.7, help="serving: offered load as a fraction of capacity")
    ap.add_argument("--min-tris", type=float, default=1e3, help="serving: smallest request (triangles, log-uniform sizes)")
    ap.add_argument("--max-tris", type=float, default=2e6, help="serving: largest request (triangles)")
    ap.add_argument("--cpu-threads", type=int, default=16, help="threads of the all-core CPU baseline (batches)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline timing")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one device copy of 256 MiB (reference point for PMC byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(meshes and patch batches shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "patches4096":
        out = bench_patches(args)
        if not args.no_plain:
            out["plain"] = bench_plain(args, "icosphere8")
        print(json.dumps(out), flush=True)
        return
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "sloppy":
        return bench_sloppy(args)
    if args.workload == "update":
        return bench_update(args)
    if args.workload == "clod":
        return bench_clod(args)
    print(json.dumps(bench_plain(args, args.workload)), flush=True)


def bench_plain(args, workload):
    """One plain simplify call per step on a device-resident mesh (meshopt_hip_simplify_device), and
    the same call through the drop-in host entry (meshopt_simplify: H2D of the inputs and D2H of the
    result inside the timed call)."""
    import torch
    import meshoptimizer_amd as M

    m, target, weights, desc = make_workload(workload)
    dev = torch.device("cuda:0")
    ib = torch.from_numpy(m.indices.astype(np.int32)).to(dev)
    vb = torch.from_numpy(m.vertices).to(dev)
    dest = torch.empty(m.index_count, dtype=torch.int32, device=dev)
    stride = m.stride
    attr_ptr = vb.data_ptr() + 12 if weights is not None else None

    def step():
        return M.simplify_device(dest.data_ptr(), ib.data_ptr(), m.index_count, vb.data_ptr(), m.vertex_count, stride,
                                 target, 1e-2, 0, attributes_ptr=attr_ptr, attribute_stride=stride if weights is not None else 0,
                                 weights=weights)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    if args.calibrate_copy:
        calibrate_copy(torch, dev)

    # bit-exactness of the measured configuration against the CPU reference (untimed)
    n_gpu, e_gpu = step()
    gpu_out = dest[:n_gpu].cpu().numpy().astype(np.uint32)
    call, passes = M.last_call_stats()

    times = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        times.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # the round kernel's own duration, from HIP events around each launch, in separate untimed
    # calls (the events add host work between launches, so they stay out of the timed calls)
    M.kernel_timing(True)
    timed_calls = max(1, min(args.steps, 5))
    for _ in range(timed_calls):
        step()
    torch.cuda.synchronize()
    x_ms, x_launches, x_ranks = M.kernel_timing(False)

    ms = elapsed / args.steps * 1e3
    value = m.triangle_count / (elapsed / args.steps) / 1e6

    # the drop-in host entry point on host arrays (staging copies included)
    attrs = m.vertices[:, 3:3 + len(weights)] if weights is not None else None
    host_t = []
    for k in range(max(3, min(args.steps, 10)) + 1):
        ts = time.perf_counter()
        if attrs is not None:
            M.simplify_with_attributes(m.indices, m.vertices, attrs, weights, None, target, 1e-2, 0)
        else:
            M.simplify(m.indices, m.vertices, target, 1e-2, 0)
        if k:
            host_t.append(time.perf_counter() - ts)
    host_ms = statistics.median(host_t) * 1e3

    A = 0 if weights is None else int((weights > 0).sum())
    ab = alg_bytes(m.triangle_count, call["vertex_count"], A, passes)

    out = {
        "metric": "input Mtri/s simplified (bit-exact vs CPU), plain meshopt_simplify",
        "value": round(value, 3),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": desc, "input_triangles": m.triangle_count, "output_triangles": n_gpu // 3,
                   "passes": len(passes), "rounds": sum(p["rounds"] for p in passes)},
        "p50_ms": round(statistics.median(times) * 1e3, 4),
        "p99_ms": round(float(np.percentile(np.array(times) * 1e3, 99)), 4),
        "host_entry": {"value": round(m.triangle_count / (host_ms * 1e-3) / 1e6, 3), "unit": "Mtri/s", "ms_per_call": round(host_ms, 4),
                       "sample": f"meshopt_simplify{'WithAttributes' if attrs is not None else ''} on host arrays, median of "
                                 f"{len(host_t)} calls (H2D of the inputs and D2H of the result inside each call)"},
        "call_roofline": {"bound": "hbm", "alg_bytes": ab, "achieved_GBs": round(ab / (ms * 1e-3) / 1e9, 2),
                          "peak_GBs": HBM_PEAK_GBS, "frac": round(ab / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
    }

    if x_launches:
        # the fixed formula's perform bytes of one call, spread over the round launches of one call
        launches_per_call = x_launches / timed_calls
        per_launch_bytes = perform_bytes(call["vertex_count"], passes) / launches_per_call
        per_launch_s = x_ms * 1e-3 / x_launches
        achieved = per_launch_bytes / per_launch_s / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": "k_round_x (commit rounds over many workgroups)",
                           "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(workload),
                           "alg_bytes_per_launch": round(per_launch_bytes), "alg_bytes_formula":
                               "perform term of the fixed formula, sum over passes of 16 C_p + 24 T_p + 35 V, per launch",
                           "launches_per_step": launches_per_call,
                           "avg_launch_us": round(per_launch_s * 1e6, 2),
                           "ranks_per_launch": round(x_ranks / x_launches, 1)}

    if not args.no_cpu:
        from oracle import oracle as O
        ro, eo = O.simplify(m.indices, m.vertices, target, 1e-2, 0, attributes=attrs, weights=weights,
                            impl="ref" if O.have_ref() else "oracle")
        out["bit_exact"] = bool(np.array_equal(ro, gpu_out) and np.float32(eo).tobytes() == np.float32(e_gpu).tobytes())
        out["cpu_baseline"] = cpu_reference(m, target, weights)
    return out


if __name__ == "__main__":
    main()
