"""GPU parity of one clusterlod level on the device (meshopt_hip_clod_level_device): the level's
boundary locks and every group's simplification, against the oracle's Sparse call per group on the
shared vertex buffer (the call demo/clusterlod.h:442-451 makes) with the locks lockBoundary
(demo/clusterlod.h:372-419) computes."""
import sys

import numpy as np
import pytest

from meshes import lock_boundary, same_bits
from oracle import oracle as O
import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W

pytestmark = pytest.mark.gpu

FLT_MAX = float(np.finfo(np.float32).max)


def announce(what):
    print(f"[case] {what}", flush=True)
    sys.stdout.flush()


def level_case(n, tiles, seed, bands=None, attributes=None):
    """A seamed grid's positions, its triangles split into tiles x tiles spatial groups (or, with
    bands, into vertical bands cut at those x fractions), in shuffled group order, the position
    remap, random caller lock/protect bits, and the expected locks: lockBoundary's own loops
    restated step by step (tests/meshes.py).  With `attributes` (a list), the grid's five
    attributes are appended to it."""
    rng = np.random.default_rng(seed)
    m = W.seamed_grid(n, seams=4, seed=seed)
    v = np.ascontiguousarray(m.vertices[:, :3])
    if attributes is not None:
        attributes.append(np.ascontiguousarray(m.vertices[:, 3:8]))
    tri = m.indices.reshape(-1, 3)
    c = v[tri].mean(axis=1)
    if bands is None:
        g = np.minimum((c[:, 1] * tiles).astype(int), tiles - 1) * tiles + np.minimum((c[:, 0] * tiles).astype(int), tiles - 1)
    else:
        g = np.searchsorted(np.array(bands), c[:, 0])
        tiles = int(np.ceil(np.sqrt(len(bands) + 1)))
    ids = rng.permutation(tiles * tiles)
    groups = [tri[g == k].reshape(-1).astype(np.uint32) for k in ids if (g == k).any()]
    _, first, inv = np.unique(v, axis=0, return_index=True, return_inverse=True)
    remap = first[inv.reshape(-1)].astype(np.uint32)
    base = ((rng.random(v.shape[0]) < 0.05) * M.SimplifyVertex_Lock + (rng.random(v.shape[0]) < 0.1) * M.SimplifyVertex_Protect).astype(np.uint8)
    locks = lock_boundary(remap, groups, base)
    return v, groups, remap, base, locks


@pytest.mark.parametrize("n,tiles,seed,opts", [(48, 3, 1, M.SimplifyErrorAbsolute | M.SimplifyPermissive),
                                               (64, 4, 2, M.SimplifyErrorAbsolute | M.SimplifyPermissive),
                                               (96, 6, 3, M.SimplifyErrorAbsolute | M.SimplifyPermissive | M.SimplifySparse),
                                               (64, 5, 4, M.SimplifyErrorAbsolute),
                                               (60, 2, 7, M.SimplifyErrorAbsolute | M.SimplifyPermissive),
                                               (96, 2, 6, M.SimplifyErrorAbsolute | M.SimplifyPermissive)])
def test_clod_level(n, tiles, seed, opts):
    """Groups of about 1000-1500 indices (hashed in LDS, simplified in the batched kernel), of about
    5400 (60x60 grid in 2x2 tiles: the default clusterlod group size, global hash slices, the
    one-workgroup-per-CU class) and, on a 96x96 grid in 2x2 tiles, of about 14000 indices
    (global hash slices, the plain path)."""
    from devmem import DeviceArray
    v, groups, remap, base, locks_expected = level_case(n, tiles, seed)
    io = np.zeros(len(groups) + 1, np.uint32)
    io[1:] = np.cumsum([grp.size for grp in groups])
    idx = np.concatenate(groups).astype(np.uint32)
    targets = np.array([grp.size // 3 // 2 * 3 for grp in groups], np.uint32)
    G = len(groups)
    announce(f"clod level grid {n} groups {G} seed {seed} opts {opts}")
    with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(v) as d_v, DeviceArray(remap) as d_remap, \
            DeviceArray(base) as d_base, DeviceArray(nbytes=v.shape[0]) as d_locks, DeviceArray(targets) as d_tg, \
            DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * G) as d_cnt, DeviceArray(nbytes=4 * G) as d_err:
        M.clod_level_device(d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, G, d_v.ptr, v.shape[0], 12, d_remap.ptr, d_base.ptr,
                            d_locks.ptr, d_tg.ptr, FLT_MAX, opts)
        locks = d_locks.to_host(np.uint8, v.shape[0])
        counts = d_cnt.to_host(np.uint32, G)
        errors = d_err.to_host(np.float32, G)
        dst = d_dst.to_host(np.uint32, idx.size)
    assert np.array_equal(locks, locks_expected), f"{int((locks != locks_expected).sum())} lock bytes differ"
    moved = 0
    for p, grp in enumerate(groups):
        ro, eo = O.simplify(grp, v, int(targets[p]), FLT_MAX, opts | M.SimplifySparse, attributes=np.zeros((v.shape[0], 0), np.float32),
                            weights=np.zeros(0, np.float32), lock=locks_expected)
        rg = dst[io[p]:io[p] + counts[p]]
        assert rg.size == ro.size and np.array_equal(rg, ro), f"group {p}: indices differ from the Sparse call"
        assert same_bits(errors[p], eo), f"group {p}: result_error {errors[p]!r} != {eo!r}"
        moved += rg.size < grp.size
    assert moved > 0


def test_clod_level_mixed_group_sizes():
    """One level whose groups span every path: about 2000 and 5000 indices (LDS hash; the 2- and
    1-per-CU classes), about 8000-27000 (global hash slices; the plain path)."""
    from devmem import DeviceArray
    v, groups, remap, base, locks_expected = level_case(96, 0, 8, bands=[0.04, 0.14, 0.29, 0.5])
    sizes = sorted(grp.size for grp in groups)
    assert sizes[0] < 3000 and 3000 < sizes[1] <= 5952 and sizes[-1] > 20000
    io = np.zeros(len(groups) + 1, np.uint32)
    io[1:] = np.cumsum([grp.size for grp in groups])
    idx = np.concatenate(groups).astype(np.uint32)
    targets = np.array([grp.size // 3 // 3 * 3 for grp in groups], np.uint32)
    G = len(groups)
    opts = M.SimplifyErrorAbsolute | M.SimplifyPermissive
    announce(f"clod level mixed group sizes {sizes}")
    with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(v) as d_v, DeviceArray(remap) as d_remap, \
            DeviceArray(base) as d_base, DeviceArray(nbytes=v.shape[0]) as d_locks, DeviceArray(targets) as d_tg, \
            DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * G) as d_cnt, DeviceArray(nbytes=4 * G) as d_err:
        M.clod_level_device(d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, G, d_v.ptr, v.shape[0], 12, d_remap.ptr, d_base.ptr,
                            d_locks.ptr, d_tg.ptr, FLT_MAX, opts)
        locks = d_locks.to_host(np.uint8, v.shape[0])
        counts = d_cnt.to_host(np.uint32, G)
        errors = d_err.to_host(np.float32, G)
        dst = d_dst.to_host(np.uint32, idx.size)
    assert np.array_equal(locks, locks_expected)
    for p, grp in enumerate(groups):
        ro, eo = O.simplify(grp, v, int(targets[p]), FLT_MAX, opts | M.SimplifySparse, attributes=np.zeros((v.shape[0], 0), np.float32),
                            weights=np.zeros(0, np.float32), lock=locks_expected)
        rg = dst[io[p]:io[p] + counts[p]]
        assert rg.size == ro.size and np.array_equal(rg, ro), f"group {p} ({grp.size} indices): indices differ from the Sparse call"
        assert same_bits(errors[p], eo), f"group {p}: result_error {errors[p]!r} != {eo!r}"


def test_clod_level_rejects_bad_input():
    """An index beyond the vertex buffer and inconsistent group offsets are reported as errors."""
    from devmem import DeviceArray
    v, groups, remap, base, _ = level_case(16, 2, 5)
    idx = np.concatenate(groups).astype(np.uint32)
    io = np.array([0, idx.size], np.uint32)
    bad = idx.copy()
    bad[4] = v.shape[0] + 2
    tg = np.array([0], np.uint32)
    announce("clod level bad input")
    with DeviceArray(bad) as d_bad, DeviceArray(np.array([0, 7], np.uint32)) as d_io_bad, DeviceArray(io) as d_io, DeviceArray(v) as d_v, \
            DeviceArray(remap) as d_remap, DeviceArray(nbytes=v.shape[0]) as d_locks, DeviceArray(tg) as d_tg, \
            DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4) as d_cnt:
        with pytest.raises(RuntimeError, match="vertex_count"):
            M.clod_level_device(d_dst.ptr, d_cnt.ptr, None, d_bad.ptr, d_io.ptr, 1, d_v.ptr, v.shape[0], 12, d_remap.ptr, None, d_locks.ptr,
                                d_tg.ptr, FLT_MAX, 0)
        with pytest.raises(RuntimeError, match="index counts"):
            M.clod_level_device(d_dst.ptr, d_cnt.ptr, None, d_bad.ptr, d_io_bad.ptr, 1, d_v.ptr, v.shape[0], 12, d_remap.ptr, None,
                                d_locks.ptr, d_tg.ptr, FLT_MAX, 0)


FALLBACK_CASES = [
    # (grid, tiles, seed, options, lock fraction, target fraction, config)
    (64, 4, 11, M.SimplifyErrorAbsolute, 0.5, 8,
     {"fallback_permissive": 1, "fallback_sloppy": 1, "error_factor_sloppy": 2.0, "error_edge_limit": 0.0}),
    (64, 4, 12, M.SimplifyErrorAbsolute | M.SimplifyPermissive, 0.5, 8,
     {"fallback_permissive": 0, "fallback_sloppy": 1, "error_factor_sloppy": 2.0, "error_edge_limit": 0.0}),
    (48, 3, 13, M.SimplifyErrorAbsolute, 0.15, 4,
     {"fallback_permissive": 1, "fallback_sloppy": 0, "error_factor_sloppy": 1.0, "error_edge_limit": 0.0}),
    (64, 5, 14, M.SimplifyErrorAbsolute, 0.4, 6,
     {"fallback_permissive": 1, "fallback_sloppy": 1, "error_factor_sloppy": 3.0, "error_edge_limit": 0.25}),
    (96, 2, 15, M.SimplifyErrorAbsolute, 0.3, 6,
     {"fallback_permissive": 1, "fallback_sloppy": 1, "error_factor_sloppy": 2.0, "error_edge_limit": 1.0}),
]


@pytest.mark.parametrize("n,tiles,seed,opts,lock_frac,div,config", FALLBACK_CASES)
def test_clod_level_fallbacks(n, tiles, seed, opts, lock_frac, div, config):
    """clodBuild's per-group simplify with its fallbacks (demo/clusterlod.h:461-518): groups that
    heavy caller locks and low targets leave above their target run again with Permissive and
    then through meshopt_simplifySloppy on their deindexed triangles; the edge-length limit caps
    every error; one group's target exceeds its size (it keeps its indices, error 0).  Every
    group against the oracle's restatement of the same steps (oracle/clod.py).  The last case's
    groups of about 14000 indices take the plain path in both the level's batch and the
    Permissive sub-batch."""
    from devmem import DeviceArray
    from oracle import clod as C
    from meshes import lock_boundary
    v, groups, remap, _, _ = level_case(n, tiles, seed)
    rng = np.random.default_rng(seed + 100)
    base = ((rng.random(v.shape[0]) < lock_frac) * M.SimplifyVertex_Lock).astype(np.uint8)
    locks_expected = lock_boundary(remap, groups, base)
    io = np.zeros(len(groups) + 1, np.uint32)
    io[1:] = np.cumsum([grp.size for grp in groups])
    idx = np.concatenate(groups).astype(np.uint32)
    targets = np.array([grp.size // 3 // div * 3 for grp in groups], np.uint32)
    targets[len(groups) // 2] = groups[len(groups) // 2].size + 3
    G = len(groups)
    announce(f"clod level fallbacks grid {n} groups {G} seed {seed} opts {opts} config {config}")
    with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(v) as d_v, DeviceArray(remap) as d_remap, \
            DeviceArray(base) as d_base, DeviceArray(nbytes=v.shape[0]) as d_locks, DeviceArray(targets) as d_tg, \
            DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * G) as d_cnt, DeviceArray(nbytes=4 * G) as d_err:
        M.clod_level_device(d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, G, d_v.ptr, v.shape[0], 12, d_remap.ptr, d_base.ptr,
                            d_locks.ptr, d_tg.ptr, FLT_MAX, opts, config)
        locks = d_locks.to_host(np.uint8, v.shape[0])
        counts = d_cnt.to_host(np.uint32, G)
        errors = d_err.to_host(np.float32, G)
        dst = d_dst.to_host(np.uint32, idx.size)
        # the same level without result_errors: the same indices
        M.clod_level_device(d_dst.ptr, d_cnt.ptr, None, d_idx.ptr, d_io.ptr, G, d_v.ptr, v.shape[0], 12, d_remap.ptr, d_base.ptr,
                            d_locks.ptr, d_tg.ptr, FLT_MAX, opts, config)
        counts2 = d_cnt.to_host(np.uint32, G)
        dst2 = d_dst.to_host(np.uint32, idx.size)
    assert np.array_equal(locks, locks_expected)
    assert np.array_equal(counts, counts2)
    plain = {"fallback_permissive": 0, "fallback_sloppy": 0, "error_factor_sloppy": 1.0, "error_edge_limit": 0.0}
    used = {"permissive": 0, "sloppy": 0, "edge": 0, "keep": 0}
    for p, grp in enumerate(groups):
        ro, eo = C.clod_simplify(grp, v, locks_expected, int(targets[p]), opts, config)
        rg = dst[io[p]:io[p] + counts[p]]
        assert rg.size == ro.size and np.array_equal(rg, ro), f"group {p} ({grp.size} indices, target {targets[p]}): indices differ"
        assert np.array_equal(dst2[io[p]:io[p] + counts2[p]], rg)
        assert same_bits(errors[p], eo), f"group {p}: result_error {errors[p]!r} != {eo!r}"
        if targets[p] > grp.size:
            used["keep"] += 1
            continue
        r0, _ = C.clod_simplify(grp, v, locks_expected, int(targets[p]), opts, plain)
        if r0.size > targets[p]:
            rp, _ = C.clod_simplify(grp, v, locks_expected, int(targets[p]), opts,
                                    dict(plain, fallback_permissive=config["fallback_permissive"]))
            used["permissive"] += int(not np.array_equal(rp, r0))
            used["sloppy"] += int(config["fallback_sloppy"] and rp.size > targets[p])
        if config["error_edge_limit"] > 0:
            _, e1 = C.clod_simplify(grp, v, locks_expected, int(targets[p]), opts, dict(config, error_edge_limit=0.0))
            used["edge"] += int(eo < e1)
    assert used["keep"] == 1
    if config["fallback_permissive"] and not (opts & M.SimplifyPermissive):
        assert used["permissive"] > 0, used
    if config["fallback_sloppy"]:
        assert used["sloppy"] > 0, used
    if config["error_edge_limit"] > 0:
        assert used["edge"] > 0, used


@pytest.mark.parametrize("n,tiles,seed,opts,config", [
    (64, 4, 21, M.SimplifyErrorAbsolute | M.SimplifyPermissive,
     {"fallback_permissive": 0, "fallback_sloppy": 1, "error_factor_sloppy": 2.0, "error_edge_limit": 0.0}),
    (64, 5, 22, M.SimplifyErrorAbsolute,
     {"fallback_permissive": 1, "fallback_sloppy": 1, "error_factor_sloppy": 2.0, "error_edge_limit": 0.5}),
    (96, 2, 23, M.SimplifyErrorAbsolute | M.SimplifyPermissive, None),
])
def test_clod_level_attributes(n, tiles, seed, opts, config):
    """One clusterlod level with the grid's five attributes (the demo's clodMesh attributes,
    demo/clusterlod.h:470-479): every group as the oracle's restatement of clodBuild's simplify()
    with simplifyWithAttributes(Sparse) returns it, the heavier locks of the first two cases
    sending groups through the fallbacks; the last case's groups of about 14000 indices take the
    plain path with attributes."""
    from devmem import DeviceArray
    from oracle import clod as C
    from meshes import lock_boundary
    holder = []
    v, groups, remap, base, _ = level_case(n, tiles, seed, attributes=holder)
    attrs = holder[0]
    weights = np.array([0.5, 0.0, 1.0, 0.25, 2.0], np.float32)
    rng = np.random.default_rng(seed + 100)
    if config is not None:
        base = (base | (rng.random(v.shape[0]) < 0.4) * M.SimplifyVertex_Lock).astype(np.uint8)
    locks_expected = lock_boundary(remap, groups, base)
    io = np.zeros(len(groups) + 1, np.uint32)
    io[1:] = np.cumsum([grp.size for grp in groups])
    idx = np.concatenate(groups).astype(np.uint32)
    targets = np.array([grp.size // 3 // (6 if config else 2) * 3 for grp in groups], np.uint32)
    G = len(groups)
    announce(f"clod level attributes grid {n} groups {G} seed {seed} opts {opts} config {config}")
    with DeviceArray(idx) as d_idx, DeviceArray(io) as d_io, DeviceArray(v) as d_v, DeviceArray(remap) as d_remap, \
            DeviceArray(base) as d_base, DeviceArray(nbytes=v.shape[0]) as d_locks, DeviceArray(targets) as d_tg, DeviceArray(attrs) as d_at, \
            DeviceArray(nbytes=idx.nbytes) as d_dst, DeviceArray(nbytes=4 * G) as d_cnt, DeviceArray(nbytes=4 * G) as d_err:
        M.clod_level_device(d_dst.ptr, d_cnt.ptr, d_err.ptr, d_idx.ptr, d_io.ptr, G, d_v.ptr, v.shape[0], 12, d_remap.ptr, d_base.ptr,
                            d_locks.ptr, d_tg.ptr, FLT_MAX, opts, config, attributes_ptr=d_at.ptr, attribute_stride=4 * attrs.shape[1],
                            weights=weights)
        locks = d_locks.to_host(np.uint8, v.shape[0])
        counts = d_cnt.to_host(np.uint32, G)
        errors = d_err.to_host(np.float32, G)
        dst = d_dst.to_host(np.uint32, idx.size)
    assert np.array_equal(locks, locks_expected)
    moved = 0
    for p, grp in enumerate(groups):
        ro, eo = C.clod_simplify(grp, v, locks_expected, int(targets[p]), opts, config, attributes=attrs, weights=weights)
        rg = dst[io[p]:io[p] + counts[p]]
        assert rg.size == ro.size and np.array_equal(rg, ro), f"group {p} ({grp.size} indices): indices differ"
        assert same_bits(errors[p], eo), f"group {p}: result_error {errors[p]!r} != {eo!r}"
        moved += rg.size < grp.size
    assert moved > 0
