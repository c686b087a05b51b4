"""Benchmark of the MI355X simplifier: input triangles per second for one plain
meshopt_simplify call on a device-resident mesh, next to the reference CPU library.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload icosphere8|grid256|seamed1024]

A step is one whole simplify call (setup, every pass, compaction of the result) on inputs
already resident in HBM; the timed region is bracketed by device synchronisation.  Prints one
JSON line.
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

# compulsory bytes of one rank visit in phase X of the commit rounds (the dominant kernel):
# work-list entry 16 + status 1 + two lock ranks 8 + two previous-round minima 16 +
# three 64-bit minima published 24 + survivor entry written 16
PHASE_X_BYTES_PER_RANK = 81


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


def measured_traffic(workload):
    """HBM bytes per k_round_x launch from the committed rocprofv3 PMC summary of this workload
    (FETCH_SIZE and WRITE_SIZE runs, calibrated against a copy of known size), or None."""
    import re
    path = os.path.join(ROOT, "profiles", f"k_round_x_{workload}_traffic.md")
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="icosphere8")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one device copy of 256 MiB (reference point for PMC byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(meshes and patch batches shard across GPUs with no exchange, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    import meshoptimizer_amd as M

    m, target, weights, desc = make_workload(args.workload)
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
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this dispatch should read 262144 KiB
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # bit-exactness of the measured configuration against the CPU reference (untimed)
    n_gpu, e_gpu = step()
    gpu_out = dest[:n_gpu].cpu().numpy().astype(np.uint32)
    call, passes = M.last_call_stats()

    times = []
    M.kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        times.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    x_ms, x_launches, x_ranks = M.kernel_timing(False)

    ms = elapsed / args.steps * 1e3
    value = m.triangle_count / (elapsed / args.steps) / 1e6

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
        "call_roofline": {"bound": "hbm", "alg_bytes": ab, "achieved_GBs": round(ab / (ms * 1e-3) / 1e9, 2),
                          "peak_GBs": HBM_PEAK_GBS, "frac": round(ab / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
    }

    if x_launches:
        per_launch_bytes = PHASE_X_BYTES_PER_RANK * x_ranks / x_launches
        per_launch_s = x_ms * 1e-3 / x_launches
        achieved = per_launch_bytes / per_launch_s / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": "k_round_x (commit rounds, phase X)",
                           "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(args.workload),
                           "launches_per_step": x_launches / args.steps,
                           "avg_launch_us": round(per_launch_s * 1e6, 2),
                           "ranks_per_launch": round(x_ranks / x_launches, 1)}

    if not args.no_cpu:
        from oracle import oracle as O
        ro, eo = O.simplify(m.indices, m.vertices, target, 1e-2, 0,
                            attributes=m.vertices[:, 3:3 + len(weights)] if weights is not None else None,
                            weights=weights, impl="ref" if O.have_ref() else "oracle")
        out["bit_exact"] = bool(np.array_equal(ro, gpu_out) and np.float32(eo).tobytes() == np.float32(e_gpu).tobytes())
        out["cpu_baseline"] = cpu_reference(m, target, weights)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
