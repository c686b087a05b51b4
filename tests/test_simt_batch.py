"""The batched kernel's device code (meshoptimizer_amd/csrc/batch_device.h) run on the CPU through
the SIMT harness in tests/simt: one fiber per GPU thread, rendezvous at every barrier and
wavefront operation, the LDS image followed by an inaccessible page, LDS filled with random
bytes before each mesh and a random thread interleaving.  Each mesh must come out exactly as the
oracle's meshopt_simplify (0 ulp); a deadlock, an access past the LDS layout or a hang fails the
test.  Runs without a GPU."""
import ctypes
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from meshes import random_grid, random_soup, same_bits
from oracle import oracle as O
import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
SIMT = os.path.join(HERE, "simt")
CSRC = os.path.join(os.path.dirname(HERE), "meshoptimizer_amd", "csrc")
OUT = os.path.join(SIMT, "_build", "libsimt_batch.so")


def build_simt(threads=256):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    # one build per process: parallel test workers must not load each other's half-written file
    final = OUT if threads == 256 else OUT.replace(".so", f"_{threads}.so")
    out = f"{final}.{os.getpid()}"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wno-unknown-pragmas",
           f"-DMHIP_BATCH_THREADS={threads}", "-I", os.path.join(SIMT, "shim"), "-I", CSRC, os.path.join(SIMT, "simt_batch.cpp"), "-o", out]
    subprocess.run(cmd, check=True)
    lib = ctypes.CDLL(out)
    os.replace(out, final)
    P, SZ = ctypes.c_void_p, ctypes.c_size_t
    lib.simt_simplify_batch.restype = ctypes.c_int
    lib.simt_simplify_batch.argtypes = [P, P, P, P, P, P, P, P, P, SZ, P, SZ, ctypes.c_float, ctypes.c_uint, ctypes.c_uint,
                                        ctypes.c_uint, ctypes.c_int, ctypes.c_uint, ctypes.c_ulonglong]
    lib.simt_simplify_batch_cursor.restype = ctypes.c_int
    lib.simt_simplify_batch_cursor.argtypes = [P, P, P, P, P, P, P, P, P, SZ, P, SZ, ctypes.c_float, ctypes.c_uint, ctypes.c_uint,
                                               ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong]
    lib.simt_retries.restype = ctypes.c_uint
    lib.simt_retries.argtypes = []
    lib.simt_set_vertex_lock.restype = None
    lib.simt_set_vertex_lock.argtypes = [P]
    lib.simt_set_attributes.restype = None
    lib.simt_set_attributes.argtypes = [P, SZ, P, SZ]
    return lib


@pytest.fixture(scope="module")
def simt():
    return build_simt()


@pytest.fixture(scope="module")
def simt_wide():
    """The same device code built for 1024-thread workgroups (16 wavefronts)."""
    return build_simt(1024)


@pytest.fixture(scope="module")
def simt_mid():
    """The same device code built for 512-thread workgroups (8 wavefronts), a width a tuned batch
    shape may give a size class."""
    return build_simt(512)


@pytest.fixture(scope="module")
def simt_narrow():
    """The same device code built for 128-thread workgroups (2 wavefronts), the width of a size
    class of small meshes at twelve workgroups per CU."""
    return build_simt(128)


def run_simt(lib, meshes, targets, err, opts, mode=3, seed=1, locks=None, caps=None, attrs=None, weights=None):
    idx, io, verts, vo, stride = M.pack_meshes(meshes)
    lk = None
    if locks is not None:
        lk = np.zeros(max(int(vo[-1]), 1), np.uint8)
        for p, l in enumerate(locks):
            lk[vo[p]:vo[p + 1]] = l
    lib.simt_set_vertex_lock(None if lk is None else lk.ctypes.data)
    at = None
    if attrs is not None:
        w = np.ascontiguousarray(weights, np.float32)
        at = np.zeros((max(int(vo[-1]), 1), w.size), np.float32)
        for p, a in enumerate(attrs):
            at[vo[p]:vo[p + 1]] = a
        lib.simt_set_attributes(at.ctypes.data, 4 * w.size, w.ctypes.data, w.size)
    tg = np.ascontiguousarray(targets, dtype=np.uint32)
    n = len(meshes)
    vmax = (int(np.diff(vo).max()) + 31) & ~31
    imax = (int(np.diff(io).max()) + 95) // 96 * 96
    if caps is not None:
        assert caps[0] >= vmax and caps[1] >= imax
        vmax, imax = caps
    dest = np.zeros(max(idx.size, 1), np.uint32)
    cnt = np.zeros(n, np.uint32)
    errs = np.zeros(n, np.float32)
    status = np.full(n, 99, np.uint32)
    rc = lib.simt_simplify_batch(dest.ctypes.data, cnt.ctypes.data, errs.ctypes.data, status.ctypes.data, None,
                                 idx.ctypes.data, io.ctypes.data, verts.ctypes.data, vo.ctypes.data, stride, tg.ctypes.data,
                                 n, float(err), int(opts), vmax, imax, 120, mode, seed)
    lib.simt_set_vertex_lock(None)
    lib.simt_set_attributes(None, 0, None, 0)
    assert rc == 0, "deadlock in the device code (see stderr)"
    return [(dest[io[p]:io[p] + cnt[p]].copy(), float(errs[p]), int(status[p])) for p in range(n)]


def check(lib, meshes, targets, err, opts, locks=None, attrs=None, weights=None, **kw):
    got = run_simt(lib, meshes, targets, err, opts, locks=locks, attrs=attrs, weights=weights, **kw)
    for p, ((ib, v), t, (r, e, st)) in enumerate(zip(meshes, targets, got)):
        assert st == 0, f"mesh {p}: status {st}"
        if attrs is None:
            ro, eo = O.simplify(ib, v, int(t), err, opts, lock=None if locks is None else locks[p])
        else:
            ro, eo = O.simplify(ib, v, int(t), err, opts, attributes=np.ascontiguousarray(attrs[p], np.float32),
                                weights=np.asarray(weights, np.float32), lock=None if locks is None else locks[p])
        assert r.size == ro.size and np.array_equal(r, ro), f"mesh {p}: indices differ from the oracle"
        assert same_bits(e, eo), f"mesh {p}: result_error {e!r} != {eo!r}"


def test_patches(simt):
    ms = W.patches(count=24, seed=4)
    check(simt, [(m.indices, m.vertices) for m in ms], [W.target_count(m.index_count, 1, 2) for m in ms], 1e-2,
          M.SimplifyLockBorder)


@pytest.mark.parametrize("seed", range(6))
def test_random_meshes(simt, seed):
    rng = np.random.default_rng(9100 + seed)
    opts = [0, M.SimplifyPermissive, M.SimplifyRegularize | M.SimplifyLockBorder, M.SimplifyErrorAbsolute,
            M.SimplifyInternalDebug, M.SimplifyRegularizeLight | M.SimplifyPermissive][seed]
    meshes, targets = [], []
    for k in range(8):
        if k % 2:
            ib, v = random_soup(rng, nv=int(rng.integers(3, 120)), ntri=int(rng.integers(1, 240)))
        else:
            ib, v = random_grid(rng, n=int(rng.integers(3, 14)))
        meshes.append((ib, v))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    check(simt, meshes, targets, float(rng.choice([1e-3, 1e-2, 1e-1, 1.0])), opts, seed=seed)


def test_candidate_overflow_runs_again(simt):
    """Separate triangles make about one candidate per index, more than the capped candidate
    arrays hold: such a mesh leaves with status 5 and runs again with full arrays."""
    rng = np.random.default_rng(77)
    ntri = 60
    v = rng.random((ntri * 3, 3), dtype=np.float32)
    ib = np.arange(ntri * 3, dtype=np.uint32)
    g = W.grid(5)  # smaller than the soup: the soup sets the layout's index cap
    meshes = [(ib, v), (g.indices, g.vertices)]
    check(simt, meshes, [ib.size // 2 // 3 * 3, W.target_count(g.index_count, 1, 2)], 1.0, 0)
    assert simt.simt_retries() >= 1


@pytest.mark.parametrize("seed", range(12))
def test_vertex_locks(simt, seed):
    """Per-vertex Lock, Protect and Priority bits (src/simplifier.cpp:498-530, 1095-1098) in the
    batched device code, with and without Permissive, against the oracle given each mesh's slice."""
    rng = np.random.default_rng(700 + seed)
    meshes, targets, locks = [], [], []
    for q in range(4):
        if q % 2:
            ib, v = random_grid(rng, n=int(rng.integers(4, 12)))
        else:
            ib, v = random_soup(rng, nv=int(rng.integers(12, 90)), ntri=int(rng.integers(8, 160)))
        v = np.ascontiguousarray(v[:, :3])
        meshes.append((ib, v))
        targets.append((ib.size // 3 * int(rng.integers(0, 3)) // 4) * 3)
        locks.append(rng.integers(0, 8, size=v.shape[0]).astype(np.uint8) * (rng.random(v.shape[0]) < 0.3))
    opts = [0, 32, 1 | 32, 4 | 32, 16][seed % 5]
    check(simt, meshes, targets, float(rng.choice([1e-2, 1e-1, 1.0])), opts, locks=locks, seed=seed)


def test_one_per_cu_layout(simt_wide):
    """The largest size class (one 1024-thread workgroup per CU, the whole 160 KB of LDS: 1984
    vertices, 5952 indices) holds groups of sixteen 128-triangle clusters, the clusterlod demo's
    default (demo/nanite.cpp:80, demo/clusterlod.h:552-566)."""
    g = W.grid(31)
    s = W.seamed_grid(28, seams=3, seed=8)
    assert g.index_count <= 5952 and s.index_count <= 5952 and s.vertex_count <= 1984
    vs = np.ascontiguousarray(s.vertices[:, :3])
    lk = ((np.arange(s.vertex_count) % 11) == 0).astype(np.uint8) * M.SimplifyVertex_Lock
    meshes = [(g.indices, np.ascontiguousarray(g.vertices[:, :3])), (s.indices, vs)]
    targets = [W.target_count(g.index_count, 1, 2), W.target_count(s.index_count, 1, 2)]
    check(simt_wide, meshes, targets, 1e-2, M.SimplifyErrorAbsolute | M.SimplifyPermissive, locks=[np.zeros(g.vertex_count, np.uint8), lk],
          caps=(1984, 5952))


@pytest.mark.parametrize("seed", range(3))
def test_wide_workgroup(simt_wide, seed):
    """With 1024 threads per workgroup every loop strides by 1024 and the sort keeps 16 wavefront
    counts; meshes of every size (patches, random soups and grids, the largest class's layout)
    come out exactly as the oracle's."""
    rng = np.random.default_rng(990 + seed)
    meshes, targets = [], []
    for k in range(5):
        if k % 2:
            ib, v = random_soup(rng, nv=int(rng.integers(3, 200)), ntri=int(rng.integers(1, 400)))
        else:
            ib, v = random_grid(rng, n=int(rng.integers(3, 30)))
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    opts = [M.SimplifyErrorAbsolute | M.SimplifyPermissive, M.SimplifyLockBorder, 0][seed]
    check(simt_wide, meshes, targets, 1e-2, opts, caps=(1984, 5952) if seed == 0 else None)


@pytest.mark.parametrize("width,seed", [(512, 0), (512, 1), (128, 0), (128, 1)])
def test_mid_workgroup(simt_mid, simt_narrow, width, seed):
    """The 512- and 128-thread builds: patches, soups and grids against the oracle, with one
    workgroup per mesh and with workgroups interleaved over one cursor."""
    lib = simt_mid if width == 512 else simt_narrow
    rng = np.random.default_rng(770 + seed + (width == 128) * 10)
    meshes, targets = [], []
    for k in range(6):
        if k % 3 == 0:
            ib, v = random_soup(rng, nv=int(rng.integers(3, 200)), ntri=int(rng.integers(1, 400)))
        else:
            ib, v = random_grid(rng, n=int(rng.integers(3, 24)), seams=k % 2)
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    opts = [M.SimplifyLockBorder, M.SimplifyErrorAbsolute | M.SimplifyPermissive][seed]
    check(lib, meshes, targets, 1e-2, opts, seed=seed)
    got = run_cursor(lib, meshes, targets, 1e-2, opts, groups=2 + seed, seed=seed)
    for p, ((ib, v), t, (rg, eg, st)) in enumerate(zip(meshes, targets, got)):
        ro, eo = O.simplify(ib, v, int(t), 1e-2, opts)
        assert st == 0 and np.array_equal(rg, ro) and same_bits(eg, eo), f"mesh {p}: cursor form differs from the oracle"


@pytest.mark.parametrize("opts", [0, M.SimplifyLockBorder | M.SimplifyPermissive])
def test_edge_values(simt, opts):
    """NaN, +-0 and denormal positions (tests/meshes.py edge_value_case) through the LDS position
    remap and box of the batched device code, each mesh against the oracle."""
    from meshes import edge_value_case
    meshes, targets = [], []
    for seed in range(20):
        ib, v, kind = edge_value_case(seed)
        rng = np.random.default_rng(seed)
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    check(simt, meshes, targets, 1e-1, opts, seed=int(opts) + 3)


def run_cursor(lib, meshes, targets, err, opts, groups, seed):
    """The largest class's wrapper (simplify_from_cursor) with `groups` workgroups interleaved over
    one shared cursor, the list in a random order; returns [(indices, error, status)]."""
    idx, io, verts, vo, stride = M.pack_meshes(meshes)
    lib.simt_set_vertex_lock(None)
    tg = np.ascontiguousarray(targets, dtype=np.uint32)
    n = len(meshes)
    ids = np.random.default_rng(seed).permutation(n).astype(np.uint32)
    vmax = (int(np.diff(vo).max()) + 31) & ~31
    imax = (int(np.diff(io).max()) + 95) // 96 * 96
    dest = np.zeros(max(idx.size, 1), np.uint32)
    cnt = np.zeros(n, np.uint32)
    errs = np.zeros(n, np.float32)
    status = np.full(n, 99, np.uint32)
    rc = lib.simt_simplify_batch_cursor(dest.ctypes.data, cnt.ctypes.data, errs.ctypes.data, status.ctypes.data, ids.ctypes.data,
                                        idx.ctypes.data, io.ctypes.data, verts.ctypes.data, vo.ctypes.data, stride, tg.ctypes.data,
                                        n, float(err), int(opts), vmax, imax, groups, 300, seed)
    assert rc != -1, "deadlock in the cursor wrapper (see stderr)"
    assert rc == 0, f"cursor wrapper: rc {rc} (-2: the shared cursor did not end at meshes + workgroups)"
    return [(dest[io[p]:io[p] + cnt[p]].copy(), float(errs[p]), int(status[p])) for p in range(n)]


@pytest.mark.parametrize("seed", range(3))
def test_cursor_wrapper(simt, seed):
    """k_simplify_patches_wide's loop (simplify_from_cursor: thread 0 takes a list position with an
    atomic, hands it over through LDS, the workgroup reads it back wave-uniform) with more meshes
    than workgroups, the workgroups interleaved in random orders over the shared cursor: every mesh
    runs exactly once, on LDS its workgroup's previous mesh left behind, and comes out exactly as
    the oracle's; a barrier that not every thread reaches is reported as a deadlock."""
    rng = np.random.default_rng(880 + seed)
    meshes, targets = [], []
    for k in range(10):
        if k % 3 == 1:
            ib, v = random_soup(rng, nv=int(rng.integers(3, 120)), ntri=int(rng.integers(1, 200)))
        else:
            ib, v = random_grid(rng, n=int(rng.integers(3, 14)))
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    opts = [M.SimplifyErrorAbsolute | M.SimplifyPermissive, M.SimplifyLockBorder, 0][seed]
    got = run_cursor(simt, meshes, targets, 1e-2, opts, groups=3 + seed, seed=seed)
    for p, ((ib, v), t, (r, e, st)) in enumerate(zip(meshes, targets, got)):
        assert st == 0, f"mesh {p}: status {st}"
        ro, eo = O.simplify(ib, v, int(t), 1e-2, opts)
        assert r.size == ro.size and np.array_equal(r, ro), f"mesh {p}: indices differ from the oracle"
        assert same_bits(e, eo), f"mesh {p}: result_error {e!r} != {eo!r}"


def test_cursor_wrapper_wide(simt_wide):
    """The same wrapper in the 1024-thread build the largest class runs: two workgroups, five
    meshes."""
    rng = np.random.default_rng(8899)
    meshes, targets = [], []
    for k in range(5):
        ib, v = random_grid(rng, n=int(rng.integers(4, 12)))
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        targets.append((ib.size // 3 * int(rng.integers(1, 4)) // 4) * 3)
    opts = M.SimplifyErrorAbsolute | M.SimplifyPermissive
    got = run_cursor(simt_wide, meshes, targets, 1e-2, opts, groups=2, seed=5)
    for p, ((ib, v), t, (r, e, st)) in enumerate(zip(meshes, targets, got)):
        assert st == 0, f"mesh {p}: status {st}"
        ro, eo = O.simplify(ib, v, int(t), 1e-2, opts)
        assert r.size == ro.size and np.array_equal(r, ro) and same_bits(e, eo), f"mesh {p}: differs from the oracle"


@pytest.mark.parametrize("sanitizer", ["undefined", "address"])
def test_sanitized_build(sanitizer):
    """The batched device code built on the host with UBSan or ASan (the harness's fibers switch
    stacks, which ASan tolerates with a warning), run on patches, soups, seamed grids and edge-value
    meshes in a child process, every mesh checked against the oracle."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    out = OUT.replace(".so", f"_{sanitizer}.{os.getpid()}.so")
    flags = ["-fsanitize=undefined", "-fno-sanitize-recover=all"] if sanitizer == "undefined" else ["-fsanitize=address"]
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wno-unknown-pragmas", *flags,
           "-DMHIP_BATCH_THREADS=256", "-I", os.path.join(SIMT, "shim"), "-I", CSRC, os.path.join(SIMT, "simt_batch.cpp"), "-o", out]
    subprocess.run(cmd, check=True)
    env = dict(os.environ)
    if sanitizer == "address":
        env["LD_PRELOAD"] = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
        env["ASAN_OPTIONS"] = "detect_leaks=0"
    try:
        res = subprocess.run([sys.executable, os.path.join(SIMT, "run_sanitized.py"), out], capture_output=True, text=True, env=env, timeout=600)
    finally:
        os.remove(out)
    assert res.returncode == 0 and "sanitized run ok" in res.stdout, res.stdout[-2000:] + res.stderr[-4000:]


@pytest.mark.parametrize("seed", range(6))
def test_attributes(simt, simt_narrow, seed):
    """Batches with attributes (the kAttr build of the device code: rescaled attributes, attribute
    quadrics per wedge, the attribute terms of the rank with seam pairs and complex wedges, the
    wedge-ordered attribute update): seamed grids with their five attributes and soups with one,
    weights with zeros among them, against the oracle's simplifyWithAttributes per mesh."""
    rng = np.random.default_rng(4400 + seed)
    lib = simt if seed % 2 == 0 else simt_narrow
    opts = [0, M.SimplifyPermissive, M.SimplifyLockBorder | M.SimplifyErrorAbsolute, M.SimplifyRegularize,
            M.SimplifyPermissive | M.SimplifyErrorAbsolute, M.SimplifyRegularizeLight][seed]
    meshes, targets, attrs, locks = [], [], [], []
    soups = seed in (1, 4)
    weights = np.array([0.5], np.float32) if soups else np.array([[1.0, 0.5, 0.0, 2.0, 0.25], [0.0, 0.0, 0.3, 0.0, 0.0],
                                                                  list(W.SEAMED_GRID_WEIGHTS)][seed % 3], np.float32)
    for k in range(5):
        if soups:
            ib, v = random_soup(rng, nv=int(rng.integers(3, 120)), ntri=int(rng.integers(1, 240)))
            a = v[:, 3:4]
        else:
            ib, v = random_grid(rng, n=int(rng.integers(4, 18)), seams=int(rng.integers(0, 4)))
            a = v[:, 3:8]
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        attrs.append(np.ascontiguousarray(a, np.float32))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
        locks.append(((rng.random(v.shape[0]) < 0.1) * M.SimplifyVertex_Protect).astype(np.uint8))
    check(lib, meshes, targets, 1e-1, opts, locks=locks if seed % 2 else None, attrs=attrs, weights=weights, seed=seed)
