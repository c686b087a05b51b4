"""GPU checks of meshopt_hip_simplify_batch_device_noalloc, the batched device entry that allocates
nothing (its scratch is a static device array): through the C ABI its results equal those of
meshopt_hip_simplify_batch_device bit for bit (index buffers, counts, result_error bits), on two calls
in a row, the second with a smaller batch; its size limits and option limits fail before any launch.
Runs last: a fresh process and one that ran every other test must both pass."""
import sys

import numpy as np
import pytest

from meshes import random_grid
import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def announce(what):
    print(f"[case] {what}", flush=True)
    sys.stdout.flush()


def mixed_batch(seed, patches, grids, wide):
    """Clusterlod-style patches (the two smallest size classes), seamed grids of 2400-2904 indices
    (the 2-per-CU class) and of 3174-5400 indices (the one-workgroup-per-CU class), shuffled."""
    rng = np.random.default_rng(seed)
    ms = W.patches(count=patches, seed=seed)
    meshes = [(m.indices, m.vertices) for m in ms]
    targets = [W.target_count(m.index_count, 1, 2) for m in ms]
    for k in range(grids + wide):
        n = int(rng.integers(21, 24)) if k < grids else int(rng.integers(24, 32))
        ib, v = random_grid(rng, n=n, seams=3)
        meshes.append((ib, v))
        targets.append((ib.size // 3 * int(rng.integers(1, 4)) // 4) * 3)
    order = rng.permutation(len(meshes))
    return [meshes[i] for i in order], [targets[i] for i in order]


def run_entry(fn, packed, targets, err, opts):
    from devmem import DeviceArray
    idx, io, verts, vo, stride = packed
    tg = np.ascontiguousarray(targets, dtype=np.uint32)
    n = tg.size
    with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(verts) as d_v, DeviceArray(vo) as d_vo, DeviceArray(tg) as d_tg, \
            DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * n) as d_cnt, DeviceArray(nbytes=4 * n) as d_err:
        fn(d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, d_v.ptr, d_vo.ptr, int(vo[-1]), stride, d_tg.ptr, n, err, opts)
        cnt = d_cnt.to_host(np.uint32, n)
        return d_dst.to_host(np.uint32, idx.size), cnt, d_err.to_host(np.float32, n)


def check_same(packed, targets, err, opts):
    io = packed[1]
    dst_a, cnt_a, err_a = run_entry(M.simplify_batch_device, packed, targets, err, opts)
    dst_b, cnt_b, err_b = run_entry(M.simplify_batch_device_noalloc, packed, targets, err, opts)
    assert np.array_equal(cnt_a, cnt_b), f"index counts differ, first mesh {int(np.argmax(cnt_a != cnt_b))}"
    assert np.array_equal(err_a.view(np.uint32), err_b.view(np.uint32)), "result_error bits differ"
    for p in range(cnt_a.size):
        a = int(io[p])
        assert np.array_equal(dst_a[a:a + cnt_a[p]], dst_b[a:a + cnt_b[p]]), f"mesh {p}: indices differ"
    assert int(cnt_a.sum()) < int(io[-1])


@pytest.mark.parametrize("opts", [M.SimplifyLockBorder, M.SimplifyErrorAbsolute | M.SimplifyPermissive])
def test_noalloc_equals_batch_device_two_calls(opts):
    """A batch over every LDS size class, then a smaller one (fewer meshes and vertices: the
    scratch keeps the first call's contents, so whatever must start at zero is cleared per call)."""
    meshes, targets = mixed_batch(77 + opts, 700, 24, 300)
    packed = M.pack_meshes(meshes)
    announce(f"noalloc call 1: {len(meshes)} meshes, {int(packed[3][-1])} vertices, opts {opts}")
    check_same(packed, targets, 1e-2, opts)
    small, small_t = mixed_batch(91 + opts, 96, 4, 6)
    packed2 = M.pack_meshes(small)
    assert len(small) < len(meshes) and packed2[3][-1] < packed[3][-1]
    announce(f"noalloc call 2: {len(small)} meshes, {int(packed2[3][-1])} vertices, opts {opts}")
    check_same(packed2, small_t, 1e-2, opts)


def test_noalloc_limits_fail_before_launch():
    """More meshes than the static host words hold, more vertices than the scratch holds, and the
    Sparse option (plain path) are refused on the host; a mesh beyond every LDS class fails the call
    (the entry does not run the plain path)."""
    from devmem import DeviceArray
    meshes, targets = mixed_batch(5, 8, 0, 0)
    idx, io, verts, vo, stride = M.pack_meshes(meshes)
    tg = np.ascontiguousarray(targets, dtype=np.uint32)
    n = tg.size
    with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(verts) as d_v, DeviceArray(vo) as d_vo, DeviceArray(tg) as d_tg, \
            DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * n) as d_cnt, DeviceArray(nbytes=4 * n) as d_err:
        args = (d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, d_v.ptr, d_vo.ptr)
        announce("noalloc limits: mesh count")
        with pytest.raises(RuntimeError, match="16384"):
            M.simplify_batch_device_noalloc(*args, int(vo[-1]), stride, d_tg.ptr, 16385, 1e-2, 0)
        announce("noalloc limits: vertex count")
        with pytest.raises(RuntimeError, match="scratch"):
            M.simplify_batch_device_noalloc(*args, 2_000_000, stride, d_tg.ptr, n, 1e-2, 0)
        announce("noalloc limits: Sparse")
        with pytest.raises(RuntimeError, match="Sparse"):
            M.simplify_batch_device_noalloc(*args, int(vo[-1]), stride, d_tg.ptr, n, 1e-2, M.SimplifySparse)
        # the call after the refusals still works
        announce("noalloc after refusals")
        M.simplify_batch_device_noalloc(*args, int(vo[-1]), stride, d_tg.ptr, n, 1e-2, 0)
    big = W.grid(34)
    assert big.index_count > 5952
    meshes.append((big.indices, big.vertices))
    targets.append(W.target_count(big.index_count, 1, 2))
    announce("noalloc: a mesh beyond every LDS class")
    with pytest.raises(RuntimeError, match="fits none"):
        run_entry(M.simplify_batch_device_noalloc, M.pack_meshes(meshes), targets, 1e-2, 0)
