"""GPU parity of the batched and serving forms at the benchmark workloads' stated shapes:
  * all 4096 clusterlod-style patches (64-512 triangles cut from a displaced 2048x2048 grid,
    LockBorder, target 50%, error 1e-2) in ONE batched call, through the device entry (offsets on
    the device) and the host entry, every patch against the oracle (index buffer, count and
    result_error bits) -- the per-group call pattern of demo/clusterlod.h:654-699 and
    demo/main.cpp:525-637;
  * the serving stream: 64 requests of log-uniform 1k-2M triangles (grids and icospheres, a third
    seamed grids with normal+UV attributes, target ratio 0.5/0.25/0.1), submitted at their Poisson
    arrival offsets to a 4-slot server with a 64 GB budget; every result against the request run
    alone through meshopt_simplify / meshopt_simplifyWithAttributes and against the oracle.
The oracle runs on a thread pool of the host's cores (each call is independent)."""
import concurrent.futures as cf
import os
import sys
import threading
import time

import numpy as np
import pytest

from meshes import same_bits
from oracle import oracle as O
import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def announce(what):
    print(f"[case] {what}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def patches4096():
    ms = W.patches(4096)
    meshes = [(m.indices, m.vertices) for m in ms]
    targets = np.array([W.target_count(m.index_count, 1, 2) for m in ms], np.uint32)
    idx, io, verts, vo, stride = M.pack_meshes(meshes)
    dest, counts, errors = O.simplify_batch(idx, io, verts, vo, targets, 1e-2, M.SimplifyLockBorder, threads=THREADS, impl="oracle")
    return meshes, targets, (idx, io, verts, vo, stride), (dest, counts, errors)


def check_all(io, expect, dst, cnt, errs):
    dest, counts, errors = expect
    bad_count = np.flatnonzero(cnt != counts)
    assert bad_count.size == 0, f"{bad_count.size} patches differ in index count, first {int(bad_count[0])}"
    bad_err = np.flatnonzero(errs.view(np.uint32) != errors.view(np.uint32))
    assert bad_err.size == 0, f"{bad_err.size} patches differ in result_error, first {int(bad_err[0])}"
    for p in range(counts.size):
        a, n = int(io[p]), int(counts[p])
        assert np.array_equal(dst[a:a + n], dest[a:a + n]), f"patch {p}: indices differ from the oracle"
    assert int(counts.sum()) < int(io[-1])


def test_patches4096_device_entry(patches4096):
    """The whole 4096-patch batch in one meshopt_hip_simplify_batch_device call (the call the
    benchmark times): the class mix, the largest-first ordering and the side-by-side class launches
    at the real shape."""
    from devmem import DeviceArray
    meshes, targets, (idx, io, verts, vo, stride), expect = patches4096
    n = len(meshes)
    announce(f"patches4096 device entry: {n} patches, {idx.size // 3} triangles")
    with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(verts) as d_v, DeviceArray(vo) as d_vo, \
            DeviceArray(targets) as d_tg, DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * n) as d_cnt, \
            DeviceArray(nbytes=4 * n) as d_err:
        for rep in range(2):  # a second call on the same device state must give the same bits
            M.simplify_batch_device(d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, d_v.ptr, d_vo.ptr, int(vo[-1]), stride,
                                    d_tg.ptr, n, 1e-2, M.SimplifyLockBorder)
            dst = d_dst.to_host(np.uint32, idx.size)
            cnt = d_cnt.to_host(np.uint32, n)
            errs = d_err.to_host(np.float32, n)
            check_all(io, expect, dst, cnt, errs)


def test_patches4096_host_entry(patches4096):
    """The same batch through meshopt_hip_simplify_batch (host arrays staged by the library)."""
    meshes, targets, (idx, io, verts, vo, stride), expect = patches4096
    announce(f"patches4096 host entry: {len(meshes)} patches")
    got = M.simplify_batch(meshes, targets, 1e-2, M.SimplifyLockBorder)
    dst = np.zeros(idx.size, np.uint32)
    cnt = np.zeros(len(meshes), np.uint32)
    errs = np.zeros(len(meshes), np.float32)
    for p, (r, e) in enumerate(got):
        dst[io[p]:io[p] + r.size] = r
        cnt[p] = r.size
        errs[p] = e
    check_all(io, expect, dst, cnt, errs)


def request_args(r):
    m = r.mesh
    attrs = m.vertices[:, 3:3 + m.attr_count] if r.weights is not None else None
    return m, attrs


def test_request_stream_in_mix_equals_alone_and_oracle():
    """The serving stream at its stated shape, requests submitted at their Poisson arrival offsets
    (200 requests/s, about 0.3 s of arrivals against about as much device work, so requests
    arrive while others run and queue behind them) to 4 slots under a 64 GB budget.  Each result
    must equal the request run alone and the oracle, bit for bit."""
    reqs = W.request_stream(64, 200.0)
    tris = [r.mesh.triangle_count for r in reqs]
    announce(f"request stream: {len(reqs)} requests, {sum(tris)} triangles, largest {max(tris)}, "
             f"{sum(1 for r in reqs if r.weights is not None)} with attributes")

    with cf.ThreadPoolExecutor(max_workers=THREADS) as ex:
        futs = []
        for r in reqs:
            m, attrs = request_args(r)
            futs.append(ex.submit(O.simplify, m.indices, m.vertices, r.target_index_count, 1e-2, 0, attributes=attrs, weights=r.weights))

        results = [None] * len(reqs)
        with M.Server(budget_bytes=64 << 30, slots=4) as srv:
            tickets = [None] * len(reqs)
            ready = threading.Event()

            def waiter():
                for k in range(len(reqs)):
                    while tickets[k] is None:
                        time.sleep(0.0002)
                    results[k] = srv.wait(tickets[k])
                ready.set()

            th = threading.Thread(target=waiter)
            th.start()
            t0 = time.perf_counter()
            try:
                for k, r in enumerate(reqs):
                    delay = r.arrival_s - reqs[0].arrival_s - (time.perf_counter() - t0)
                    if delay > 0:
                        time.sleep(delay)
                    m, attrs = request_args(r)
                    tickets[k] = srv.submit(m.indices, m.vertices, r.target_index_count, 1e-2, 0, attributes=attrs, weights=r.weights)
            finally:
                ready.wait()
                th.join()
        oracle = [f.result() for f in futs]

    spans = [(i["start_ms"], i["finish_ms"], i["slot"]) for _, _, i in results]
    lat = np.array([i["finish_ms"] - i["submit_ms"] for _, _, i in results])
    announce(f"served: p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms")
    for k, r in enumerate(reqs):
        m, attrs = request_args(r)
        ri, re, _ = results[k]
        if attrs is not None:
            ai, ae = M.simplify_with_attributes(m.indices, m.vertices, attrs, r.weights, None, r.target_index_count, 1e-2, 0)
        else:
            ai, ae = M.simplify(m.indices, m.vertices, r.target_index_count, 1e-2, 0)
        assert np.array_equal(ri, ai) and same_bits(re, ae), f"request {k} ({m.name}): mix differs from the request alone"
        oi, oe = oracle[k]
        assert np.array_equal(ri, oi) and same_bits(re, oe), f"request {k} ({m.name}): differs from the oracle"
    overlap = any(a[0] < b[1] and b[0] < a[1] and a[2] != b[2] for x, a in enumerate(spans) for b in spans[x + 1:])
    assert overlap, "no two requests overlapped in time"
