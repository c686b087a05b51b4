"""Autotuner of the batched kernel's launch shape.

  python -m meshoptimizer_amd.tune [--patches 4096] [--seed 1] [--reps 20] [--out PATH]

The batched entry points (meshopt_hip_simplify_batch*, the clusterlod level) sort meshes into four
LDS size classes.  A class's shape is how many of its workgroups share one CU (which sets its LDS
share and with it the largest mesh it holds), the threads per workgroup (128, 256, 512 or 1024) and
whether its workgroups take one mesh each or loop over the class's list through a cursor.  The
library's default shape follows from the LDS budget alone; this tuner measures candidate shapes on a
batch with the caller's mesh-size histogram (by default the clusterlod-style workload: patches of
64-512 triangles cut from a displaced grid, LockBorder, target 50%, error 1e-2) and writes the
fastest one to the table the library reads at its first batched call (batch_shape.txt beside the
library, or the file MESHOPT_HIP_BATCH_SHAPE names).

Every candidate must give the same result as the default shape, mesh for mesh and bit for bit: a
mesh's result does not depend on the class that runs it.  A candidate whose classes cover fewer
mesh sizes than the default's (so that some mesh would fall back to the plain path) is skipped.
Timing is the batched call's own span measured with HIP events on the library's stream
(meshopt_hip_batch_kernel_timing), best of three runs of `reps` calls.
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "batch_shape.txt")

# per class: (workgroups per CU, threads, cursor form) options; classes must not grow in workgroups
# per CU, and the wavefronts per SIMD stay within each width's register budget (6 at 128, 256 and
# 512 threads, 4 at 1024)
CLASS_OPTIONS = (
    [(12, 128, 0), (10, 128, 0), (9, 128, 0), (8, 128, 0), (6, 256, 0), (5, 256, 0)],
    [(9, 128, 0), (8, 128, 0), (7, 128, 0), (6, 128, 0), (6, 256, 0), (5, 256, 0), (4, 256, 0), (3, 512, 0)],
    [(6, 256, 0), (5, 256, 0), (4, 256, 0), (3, 256, 0), (2, 256, 0), (2, 256, 1), (2, 512, 0)],
    [(1, 1024, 1)],
)


def candidate_shapes():
    out = []
    for combo in itertools.product(*CLASS_OPTIONS):
        per_cu = [c[0] for c in combo]
        if any(per_cu[k] > per_cu[k - 1] for k in range(1, 4)):
            continue
        out.append(tuple(zip(*combo)))  # (per_cu, threads, cursor)
    return out


def write_table(path, shape, meta):
    per_cu, threads, cursor = shape
    with open(path, "w") as f:
        f.write("# launch shape of the batched kernel's four size classes, measured by meshoptimizer_amd/tune.py\n")
        for k, v in meta.items():
            f.write(f"# {k}: {v}\n")
        f.write("per_cu " + " ".join(str(x) for x in per_cu) + "\n")
        f.write("threads " + " ".join(str(x) for x in threads) + "\n")
        f.write("cursor " + " ".join(str(x) for x in cursor) + "\n")


def tune(meshes, targets, target_error, options, reps=20, log=print):
    """Measures every candidate shape on the batch; returns (best shape, results).  Leaves the
    library on the best shape."""
    import torch
    import meshoptimizer_amd as M

    idx, io, verts, vo, stride = M.pack_meshes(meshes)
    tg = np.ascontiguousarray(targets, dtype=np.uint32)
    n = tg.size
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to(dev)
    d_idx, d_io, d_vo, d_tg = t(idx), t(io), t(vo), t(tg)
    d_v = torch.from_numpy(verts).to(dev)
    d_dst = torch.empty(max(idx.size, 1), dtype=torch.int32, device=dev)
    d_cnt = torch.empty(n, dtype=torch.int32, device=dev)
    d_err = torch.empty(n, dtype=torch.float32, device=dev)

    def call():
        M.simplify_batch_device(d_dst.data_ptr(), d_cnt.data_ptr(), d_err.data_ptr(), d_idx.data_ptr(), d_io.data_ptr(), d_v.data_ptr(),
                                d_vo.data_ptr(), int(vo[-1]), stride, d_tg.data_ptr(), n, target_error, options)

    def result():
        torch.cuda.synchronize()
        cnt = d_cnt.cpu().numpy().view(np.uint32).copy()
        dst = d_dst.cpu().numpy().view(np.uint32)
        return cnt, d_err.cpu().numpy().view(np.uint32).copy(), [dst[io[p]:io[p] + cnt[p]].copy() for p in range(n)]

    def measure():
        best = None
        for _ in range(3):
            M.kernel_timing(True)
            M.batch_kernel_timing()
            for _ in range(reps):
                call()
            torch.cuda.synchronize()
            ms, launches, _ = M.batch_kernel_timing()
            M.kernel_timing(False)
            per = ms / max(launches, 1)
            best = per if best is None else min(best, per)
        return best

    M.batch_set_shape()
    default = M.batch_get_shape()
    for _ in range(3):
        call()
    ref = result()
    results = []
    for shape in candidate_shapes():
        per_cu, threads, cursor = shape
        try:
            M.batch_set_shape(per_cu, threads, cursor)
            caps = M.batch_get_shape()
        except (ValueError, RuntimeError) as e:
            log(f"skip {shape}: {e}")
            continue
        if caps["index_caps"][2] == 0 or caps["index_caps"][3] < default["index_caps"][3] or caps["vertex_caps"][3] < default["vertex_caps"][3]:
            log(f"skip {shape}: its classes hold fewer mesh sizes than the default's")
            continue
        for _ in range(2):
            call()
        got = result()
        same = np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]) and all(np.array_equal(a, b) for a, b in zip(got[2], ref[2]))
        if not same:
            raise RuntimeError(f"shape {shape} changed a result: the batched kernel is wrong in that shape")
        ms = measure()
        results.append({"per_cu": list(per_cu), "threads": list(threads), "cursor": list(cursor), "ms": ms,
                        "index_caps": caps["index_caps"], "vertex_caps": caps["vertex_caps"]})
        log(f"{list(per_cu)} {list(threads)} {list(cursor)}: {ms:.4f} ms")
    results.sort(key=lambda r: r["ms"])
    best = results[0]
    M.batch_set_shape(best["per_cu"], best["threads"], best["cursor"])
    return best, results


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--patches", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--out", default=TABLE, help="table to write (default: batch_shape.txt beside the library)")
    args = ap.parse_args(argv)

    sys.path.insert(0, os.path.dirname(HERE))
    import meshoptimizer_amd as M
    from meshoptimizer_amd import workloads as W

    ms = W.patches(count=args.patches, seed=args.seed)
    meshes = [(m.indices, m.vertices) for m in ms]
    targets = [W.target_count(m.index_count, 1, 2) for m in ms]
    t0 = time.time()
    best, results = tune(meshes, targets, 1e-2, M.SimplifyLockBorder, reps=args.reps, log=lambda s: print(s, flush=True))
    default = next(r for r in results if (r["per_cu"], r["threads"], r["cursor"]) == ([6, 5, 2, 1], [256, 256, 256, 1024], [0, 0, 0, 1]))
    meta = {"workload": f"{args.patches} clusterlod-style patches of 64-512 triangles (seed {args.seed}), LockBorder, target 50%, error 1e-2",
            "default_ms": f"{default['ms']:.4f}", "tuned_ms": f"{best['ms']:.4f}", "gain": f"{default['ms'] / best['ms']:.4f}",
            "candidates": len(results)}
    write_table(args.out, (best["per_cu"], best["threads"], best["cursor"]), meta)
    summary = {"best": best, "default_ms": default["ms"], "gain": default["ms"] / best["ms"], "candidates": len(results),
               "seconds": round(time.time() - t0, 1), "table": args.out}
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
