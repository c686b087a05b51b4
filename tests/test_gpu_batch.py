"""GPU parity of the batched entry point (meshopt_hip_simplify_batch): every mesh of a batch gets
exactly what the oracle's meshopt_simplify returns for it alone (index buffer, index count and
result_error bits)."""
import sys

import numpy as np
import pytest

from meshes import edge_value_case, random_grid, random_soup, same_bits, scattered_components
from oracle import oracle as O
import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def announce(what):
    print(f"[case] {what}", flush=True)
    sys.stdout.flush()


def check_batch(meshes, targets, err, opts, locks=None):
    got = M.simplify_batch(meshes, targets, err, opts, locks=locks)
    assert len(got) == len(meshes)
    for p, ((ib, v), t, (rg, eg)) in enumerate(zip(meshes, targets, got)):
        ro, eo = O.simplify(ib, v, int(t), err, opts, lock=None if locks is None else locks[p])
        assert rg.size == ro.size, f"mesh {p}: index count {rg.size} != oracle {ro.size}"
        assert np.array_equal(rg, ro), f"mesh {p}: first difference at {int(np.argmax(rg != ro))}"
        assert same_bits(eg, eo), f"mesh {p}: result_error {eg!r} != oracle {eo!r}"
    return got


def patch_meshes(count, seed=4):
    ms = W.patches(count=count, seed=seed)
    return [(m.indices, m.vertices) for m in ms], [W.target_count(m.index_count, 1, 2) for m in ms]


def test_clusterlod_patches():
    """The batched workload's own shape: grid patches of 64-512 triangles, LockBorder, 50%."""
    announce("patches count=96 LockBorder target 50% err 1e-2")
    meshes, targets = patch_meshes(96)
    check_batch(meshes, targets, 1e-2, M.SimplifyLockBorder)


@pytest.mark.parametrize("opts", [0, M.SimplifyErrorAbsolute, M.SimplifyRegularize, M.SimplifyPermissive,
                                  M.SimplifyRegularizeLight, M.SimplifyLockBorder | M.SimplifyInternalDebug])
def test_patches_options(opts):
    announce(f"patches count=32 opts={opts}")
    meshes, targets = patch_meshes(32, seed=11)
    check_batch(meshes, targets, 5e-2, opts)


OPTION_SETS = [0, 1, 4, 16, 32, 64, 1 | 32, 32 | 16, 1 | 4, 1 << 30]


@pytest.mark.parametrize("seed", range(12))
def test_random_batches(seed):
    """Mixed random soups (repeated positions, degenerate triangles) and seamed grids in one batch."""
    rng = np.random.default_rng(7000 + seed)
    opts = OPTION_SETS[seed % len(OPTION_SETS)]
    err = float(rng.choice([1e-3, 1e-2, 1e-1, 1.0]))
    meshes, targets = [], []
    for k in range(int(rng.integers(4, 24))):
        if k % 2:
            ib, v = random_soup(rng, nv=int(rng.integers(3, 150)), ntri=int(rng.integers(1, 300)))
        else:
            ib, v = random_grid(rng, n=int(rng.integers(3, 16)))
        meshes.append((ib, v))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    announce(f"random batch seed={seed} meshes={len(meshes)} opts={opts} err={err}")
    check_batch(meshes, targets, err, opts)


def test_oversized_mesh_takes_plain_path():
    """A mesh too large for one workgroup's LDS is simplified by the plain path inside the same
    call; the small meshes around it still run in the batched kernel."""
    big = W.grid(160)
    meshes, targets = patch_meshes(12, seed=21)
    meshes.insert(5, (big.indices, big.vertices))
    targets.insert(5, W.target_count(big.index_count, 1, 4))
    announce(f"batch with one mesh of {big.index_count // 3} triangles")
    check_batch(meshes, targets, 1e-2, M.SimplifyLockBorder)


class launch_shape:
    """Runs a block under the given launch shape (none: the built-in default), then restores the
    shape that was in use (the shipped table, unless a test set another)."""

    def __init__(self, *shape):
        self.shape = shape

    def __enter__(self):
        self.saved = M.batch_get_shape()
        M.batch_set_shape(*self.shape)
        return M.batch_get_shape()

    def __exit__(self, *exc):
        M.batch_set_shape(self.saved["per_cu"], self.saved["threads"], self.saved["cursor"])
        return False


def check_batch_device(meshes, targets, err, opts):
    """The device entry (offsets on the device only) against the oracle, mesh by mesh."""
    from devmem import DeviceArray
    idx, io, verts, vo, stride = M.pack_meshes(meshes)
    tg = np.ascontiguousarray(targets, dtype=np.uint32)
    n = len(meshes)
    with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(verts) as d_v, DeviceArray(vo) as d_vo, DeviceArray(tg) as d_tg, \
            DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * n) as d_cnt, DeviceArray(nbytes=4 * n) as d_err:
        M.simplify_batch_device(d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, d_v.ptr, d_vo.ptr, int(vo[-1]), stride, d_tg.ptr, n, err, opts)
        dst = d_dst.to_host(np.uint32, idx.size)
        cnt = d_cnt.to_host(np.uint32, n)
        errs = d_err.to_host(np.float32, n)
    for p, ((ib, v), t) in enumerate(zip(meshes, targets)):
        ro, eo = O.simplify(ib, v, int(t), err, opts)
        rg = dst[io[p]:io[p] + cnt[p]]
        assert rg.size == ro.size and np.array_equal(rg, ro), f"mesh {p}: indices differ from the oracle"
        assert same_bits(errs[p], eo), f"mesh {p}: result_error {errs[p]!r} != {eo!r}"


@pytest.mark.parametrize("entry", ["host", "device"])
@pytest.mark.parametrize("opts", [M.SimplifyLockBorder, M.SimplifyErrorAbsolute | M.SimplifyPermissive])
def test_size_classes(opts, entry):
    """Meshes for every LDS size class of the built-in default shape in one call: patches (6 and
    5 workgroups per CU), seamed grids of 2400-2904 indices (2 per CU) and 300 of 4056-5400
    indices (one workgroup per CU with the whole LDS, taking meshes from a cursor: more of them
    than CUs), and one of 6534 indices beyond every class (the plain path).  The host entry knows
    the sizes and launches the largest class beside the others; the device entry launches it
    after the status read-back."""
    rng = np.random.default_rng(4242 + opts)
    meshes, targets = patch_meshes(40, seed=31)
    for k in range(320):
        n = int(rng.integers(21, 24)) if k < 20 else int(rng.integers(27, 32))
        ib, v = random_grid(rng, n=n, seams=3)
        meshes.append((ib, v))
        targets.append((ib.size // 3 * int(rng.integers(1, 4)) // 4) * 3)
    big = W.grid(34)
    assert big.index_count > 5952
    meshes.append((big.indices, big.vertices))
    targets.append(W.target_count(big.index_count, 1, 2))
    order = rng.permutation(len(meshes))
    meshes = [meshes[i] for i in order]
    targets = [targets[i] for i in order]
    announce(f"size classes: {len(meshes)} meshes, largest {max(m[0].size for m in meshes)} indices, opts {opts}, {entry} entry")
    with launch_shape() as shape:
        caps = shape["index_caps"]
        assert 2904 <= caps[2] < 4056 and caps[3] >= 5400, caps
        if entry == "host":
            check_batch(meshes, targets, 1e-2, opts)
        else:
            check_batch_device(meshes, targets, 1e-2, opts)


def test_sparse_batch():
    rng = np.random.default_rng(99)
    meshes, targets = [], []
    for _ in range(6):
        ib, v = random_soup(rng, nv=200, ntri=60)
        meshes.append((ib, v))
        targets.append(ib.size // 2 // 3 * 3)
    announce("sparse batch")
    check_batch(meshes, targets, 1e-1, M.SimplifySparse)


def test_empty_and_trivial_meshes():
    announce("empty batch, empty mesh, target above the count")
    assert M.simplify_batch([], [], 1e-2, 0) == []
    m = W.grid(6)
    empty = (np.zeros(0, np.uint32), np.zeros((0, 3), np.float32))
    meshes = [empty, (m.indices, m.vertices), (m.indices, m.vertices), empty]
    targets = [0, m.index_count, 0, 0]
    check_batch(meshes, targets, 1e-2, 0)


def test_batch_beyond_fused_ordering():
    """A batch of more meshes than the one-launch classification takes (16384): the classify and
    order kernels run apart; every mesh still matches the oracle."""
    rng = np.random.default_rng(4242)
    base = [W.grid(n, seed=s) for n, s in ((4, 1), (5, 2), (6, 3), (9, 4), (14, 5), (24, 6))]
    picks = rng.integers(0, len(base), 16500)
    meshes = [(base[k].indices, base[k].vertices) for k in picks]
    targets = [int(rng.integers(0, base[k].index_count // 3 + 1)) * 3 for k in picks]
    announce(f"batch of {len(meshes)} small grids, random targets")
    check_batch(meshes, targets, 5e-2, 0)


def test_bad_index_raises():
    announce("index out of range")
    m = W.grid(5)
    ib = m.indices.copy()
    ib[7] = m.vertex_count + 3
    with pytest.raises(RuntimeError, match="vertex_count"):
        M.simplify_batch([(m.indices, m.vertices), (ib, m.vertices)], [0, 0], 1e-2, 0)


def test_prune_batch():
    """Prune batches run mesh by mesh on the plain path inside the call, bit-exact per mesh, on
    meshes of many components at very different scales (Prune removes the small ones)."""
    rng = np.random.default_rng(31)
    meshes, targets = [], []
    for k in range(5):
        ib, v = scattered_components(rng, int(rng.integers(4, 16)))
        meshes.append((ib, v))
        targets.append(0 if k % 2 else (ib.size // 3 // 2) * 3)
    announce("prune batch")
    check_batch(meshes, targets, 5e-2, M.SimplifyPrune)


@pytest.mark.parametrize("seed", range(10))
def test_vertex_locks(seed):
    """Per-vertex Lock, Protect and Priority bits in the batched kernel (one byte per packed vertex),
    with and without Permissive and ErrorAbsolute, against the oracle given each mesh's slice."""
    rng = np.random.default_rng(7100 + seed)
    meshes, targets, locks = [], [], []
    for q in range(int(rng.integers(4, 24))):
        if q % 2:
            ib, v = random_grid(rng, n=int(rng.integers(4, 14)))
        else:
            ib, v = random_soup(rng, nv=int(rng.integers(12, 150)), ntri=int(rng.integers(8, 250)))
        v = np.ascontiguousarray(v[:, :3])
        meshes.append((ib, v))
        targets.append((ib.size // 3 * int(rng.integers(0, 3)) // 4) * 3)
        locks.append((rng.integers(0, 8, size=v.shape[0]) * (rng.random(v.shape[0]) < 0.3)).astype(np.uint8))
    opts = [0, M.SimplifyPermissive, M.SimplifyLockBorder | M.SimplifyPermissive, M.SimplifyErrorAbsolute | M.SimplifyPermissive,
            M.SimplifyRegularize][seed % 5]
    err = float(rng.choice([1e-2, 1e-1, 1.0]))
    announce(f"batch vertex locks seed={seed} meshes={len(meshes)} opts={opts} err={err}")
    check_batch(meshes, targets, err, opts, locks=locks)


def test_clusterlod_level_groups():
    """One clusterlod level in one batched call: the triangles of a grid split into spatial groups,
    vertices used by two groups locked (lockBoundary, demo/clusterlod.h:372-419), every group
    simplified to half with ErrorAbsolute | Permissive and no error limit (demo/clusterlod.h:442-451).
    Each group is packed with its vertices in first-use order, which is what Sparse makes of it, so
    every result must equal the oracle's Sparse call on the shared vertex buffer, mapped back."""
    m = W.grid(64)
    v = np.ascontiguousarray(m.vertices[:, :3])
    tri = m.indices.reshape(-1, 3)
    centre = v[tri].mean(axis=1)
    tiles = 4
    g = np.minimum((centre[:, 1] * tiles).astype(int), tiles - 1) * tiles + np.minimum((centre[:, 0] * tiles).astype(int), tiles - 1)
    groups = [tri[g == k].reshape(-1).astype(np.uint32) for k in range(tiles * tiles) if (g == k).any()]
    _, first, inv = np.unique(v, axis=0, return_index=True, return_inverse=True)
    remap = first[inv.reshape(-1)]
    uses = np.zeros(v.shape[0], np.int32)
    for grp in groups:
        uses[np.unique(remap[grp])] += 1
    locks = (uses >= 2).astype(np.uint8)[remap]
    flt_max = float(np.finfo(np.float32).max)

    packed, targets, plocks, orders = [], [], [], []
    for grp in groups:
        _, at = np.unique(grp, return_index=True)
        order = grp[np.sort(at)]
        local = np.zeros(v.shape[0], np.uint32)
        local[order] = np.arange(order.size, dtype=np.uint32)
        packed.append((local[grp], np.ascontiguousarray(v[order])))
        plocks.append(locks[order])
        orders.append(order)
        targets.append(grp.size // 3 // 2 * 3)
    announce(f"clusterlod level: {len(groups)} groups, {int(locks.sum())} locked vertices")
    opts = M.SimplifyErrorAbsolute | M.SimplifyPermissive
    got = M.simplify_batch(packed, targets, flt_max, opts, locks=plocks)
    for p, (grp, t, order, (rg, eg)) in enumerate(zip(groups, targets, orders, got)):
        ro, eo = O.simplify(grp, v, t, flt_max, opts | M.SimplifySparse, attributes=np.zeros((v.shape[0], 0), np.float32),
                            weights=np.zeros(0, np.float32), lock=locks)
        assert np.array_equal(order[rg], ro), f"group {p}: indices differ from the Sparse call"
        assert same_bits(eg, eo), f"group {p}: result_error {eg!r} != {eo!r}"
        assert rg.size < grp.size


@pytest.mark.parametrize("opts", [0, M.SimplifyLockBorder, M.SimplifyErrorAbsolute | M.SimplifyPermissive, M.SimplifyRegularize])
def test_edge_value_batch(opts):
    """NaN, +-0 and denormal positions in the batched kernel's LDS remap and box (one batch of 40
    meshes, every edge-value kind), each mesh against the oracle alone."""
    meshes, targets = [], []
    for seed in range(40):
        ib, v, kind = edge_value_case(seed)
        rng = np.random.default_rng(seed)
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    announce(f"edge value batch opts={opts}")
    check_batch(meshes, targets, 1e-1, opts)


@pytest.mark.parametrize("shape", [
    ([5, 4, 2, 1], [256, 256, 512, 512], [0, 0, 1, 1]),
    ([3, 3, 2, 1], [512, 512, 256, 1024], [0, 0, 1, 1]),
    ([6, 5, 3, 1], [256, 256, 256, 1024], [1, 0, 0, 1]),
    ([12, 9, 5, 1], [128, 128, 256, 1024], [0, 0, 0, 1]),
    ([10, 6, 3, 1], [128, 256, 256, 1024], [1, 0, 0, 1]),
], ids=["mid-widths", "512-first", "cursor-first", "narrow", "narrow-cursor"])
def test_batch_shapes(shape):
    """Launch shapes a tuned table may hold (other workgroups per CU, 128- and 512-thread classes,
    cursor forms): every mesh of a mix over all size classes, through the device entry, against the oracle;
    the shape in use before is restored afterwards."""
    rng = np.random.default_rng(515)
    meshes, targets = patch_meshes(48, seed=33)
    for k in range(40):
        n = int(rng.integers(21, 24)) if k < 20 else int(rng.integers(24, 32))
        ib, v = random_grid(rng, n=n, seams=3)
        meshes.append((ib, v))
        targets.append((ib.size // 3 * int(rng.integers(1, 4)) // 4) * 3)
    order = rng.permutation(len(meshes))
    meshes = [meshes[i] for i in order]
    targets = [targets[i] for i in order]
    with launch_shape(*shape) as caps:
        announce(f"batch shape {shape}: caps {caps['index_caps']}")
        check_batch_device(meshes, targets, 1e-2, M.SimplifyLockBorder)


def check_batch_attributes(meshes, attrs, weights, targets, err, opts, locks=None, entry="host"):
    """The batched entries with attributes against the oracle's simplifyWithAttributes per mesh."""
    if entry == "host":
        got = M.simplify_batch(meshes, targets, err, opts, locks=locks, attributes=attrs, weights=weights)
        res = [(rg, np.float32(eg)) for rg, eg in got]
    else:
        from devmem import DeviceArray
        idx, io, verts, vo, stride = M.pack_meshes(meshes)
        w = np.ascontiguousarray(weights, np.float32)
        at = np.zeros((max(int(vo[-1]), 1), w.size), np.float32)
        for p, a in enumerate(attrs):
            at[vo[p]:vo[p + 1]] = a
        lk = np.zeros(max(int(vo[-1]), 1), np.uint8)
        if locks is not None:
            for p, l in enumerate(locks):
                lk[vo[p]:vo[p + 1]] = l
        tg = np.ascontiguousarray(targets, dtype=np.uint32)
        n = len(meshes)
        with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(verts) as d_v, DeviceArray(vo) as d_vo, DeviceArray(tg) as d_tg, \
                DeviceArray(at) as d_at, DeviceArray(lk) as d_lk, DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * n) as d_cnt, \
                DeviceArray(nbytes=4 * n) as d_err:
            M.simplify_batch_with_attributes_device(d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, d_v.ptr, d_vo.ptr, d_at.ptr, 4 * w.size, w,
                                                    int(vo[-1]), stride, d_tg.ptr, n, err, opts, None if locks is None else d_lk.ptr)
            dst = d_dst.to_host(np.uint32, idx.size)
            cnt = d_cnt.to_host(np.uint32, n)
            errs = d_err.to_host(np.float32, n)
        res = [(dst[io[p]:io[p] + cnt[p]], errs[p]) for p in range(n)]
    for p, ((ib, v), t, (rg, eg)) in enumerate(zip(meshes, targets, res)):
        ro, eo = O.simplify(ib, v, int(t), err, opts, attributes=np.ascontiguousarray(attrs[p], np.float32), weights=np.asarray(weights, np.float32),
                            lock=None if locks is None else locks[p])
        assert rg.size == ro.size and np.array_equal(rg, ro), f"mesh {p} ({ib.size} indices): indices differ from the oracle"
        assert same_bits(eg, eo), f"mesh {p}: result_error {eg!r} != {eo!r}"


ATTRIBUTE_CASES = [
    (0, list(W.SEAMED_GRID_WEIGHTS)),
    (M.SimplifyPermissive, [1.0, 0.5, 0.0, 2.0, 0.25]),
    (M.SimplifyLockBorder | M.SimplifyErrorAbsolute, [0.0, 0.0, 0.3, 0.0, 0.0]),
    (M.SimplifyRegularize | M.SimplifyPermissive, list(W.SEAMED_GRID_WEIGHTS)),
]


@pytest.mark.parametrize("entry", ["host", "device"])
@pytest.mark.parametrize("case", range(len(ATTRIBUTE_CASES)))
def test_batch_attributes(case, entry):
    """Batches with attributes (the attribute build of the batched kernel: rescaled attributes and
    attribute quadrics in global memory, the rank's attribute terms with seam pairs and complex
    wedges, the wedge-ordered attribute update): seamed grids with their five attributes across
    every size class, caller protect bits, and one mesh beyond every class (the plain path with
    attributes), against the oracle's simplifyWithAttributes mesh by mesh."""
    opts, weights = ATTRIBUTE_CASES[case]
    rng = np.random.default_rng(6600 + case)
    meshes, attrs, targets, locks = [], [], [], []
    for k in range(48):
        n = int(rng.integers(4, 16)) if k < 36 else int(rng.integers(16, 30))
        ib, v = random_grid(rng, n=n, seams=int(rng.integers(0, 4)))
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        attrs.append(np.ascontiguousarray(v[:, 3:8]))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
        locks.append(((rng.random(v.shape[0]) < 0.1) * M.SimplifyVertex_Protect).astype(np.uint8))
    big = W.seamed_grid(44, seams=5, seed=case)
    assert big.index_count > 5952
    meshes.append((big.indices, np.ascontiguousarray(big.vertices[:, :3])))
    attrs.append(np.ascontiguousarray(big.vertices[:, 3:8]))
    targets.append(W.target_count(big.index_count, 1, 3))
    locks.append(np.zeros(big.vertices.shape[0], np.uint8))
    announce(f"batch attributes case {case} entry {entry} opts {opts} weights {weights}")
    check_batch_attributes(meshes, attrs, np.array(weights, np.float32), targets, 1e-1, opts, locks=locks, entry=entry)
