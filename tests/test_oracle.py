"""CPU checks of the oracle (the C restatement): known answers of the reference's own tests,
the compiled reference library where it is built, the round-based commit, closed forms."""
import numpy as np
import pytest

import kat
from meshes import random_grid, random_soup, same_bits, update_case
from oracle import oracle as O
from meshoptimizer_amd import workloads as W

CASES = kat.load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
@pytest.mark.parametrize("rounds", [False, True])
def test_known_answers(case, rounds):
    a = kat.arrays(case)
    attrs = None if case["call"] == "simplify" else a["attributes"]
    r, e = O.simplify(a["indices"], a["positions"], case["target_index_count"], case["target_error"], case["options"],
                      attributes=attrs, weights=a["weights"], lock=a["lock"], rounds=rounds)
    kat.check(case, r, e)


def _ref_or_skip():
    if not O.have_ref():
        pytest.skip("oracle/_ref not built (reference sources absent)")


OPTION_SETS = [0, 1, 2, 4, 16, 32, 64, 1 | 32, 2 | 8, 8, 1 | 8, 32 | 16]


@pytest.mark.parametrize("seed", range(40))
def test_oracle_matches_reference_random(seed):
    _ref_or_skip()
    rng = np.random.default_rng(seed)
    opts = OPTION_SETS[seed % len(OPTION_SETS)]
    if seed % 2:
        ib, v = random_soup(rng, nv=int(rng.integers(20, 120)), ntri=int(rng.integers(10, 300)))
    else:
        ib, v = random_grid(rng, n=int(rng.integers(5, 16)))
    target = (ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3
    err = float(rng.choice([1e-3, 1e-2, 1e-1, 1.0]))
    use_attr = seed % 3 == 0
    attrs = v[:, 3:4] if use_attr else None
    weights = np.array([0.5], np.float32) if use_attr else None
    lock = None
    if seed % 5 == 0:
        lock = rng.integers(0, 8, size=v.shape[0]).astype(np.uint8)
        attrs = v[:, 3:4]
        weights = np.array([0.25], np.float32)
    r1, e1 = O.simplify(ib, v, target, err, opts, attributes=attrs, weights=weights, lock=lock, impl="ref")
    r2, e2 = O.simplify(ib, v, target, err, opts, attributes=attrs, weights=weights, lock=lock)
    r3, e3 = O.simplify(ib, v, target, err, opts, attributes=attrs, weights=weights, lock=lock, rounds=True, window=7)
    assert r1.tolist() == r2.tolist() and same_bits(e1, e2)
    assert r1.tolist() == r3.tolist() and same_bits(e1, e3)


@pytest.mark.parametrize("name", ["grid64", "icosphere4", "seamed64", "patches"])
def test_oracle_matches_reference_workloads(name):
    _ref_or_skip()
    if name == "grid64":
        meshes, ratio, attrs = [W.grid(64)], (1, 4), False
    elif name == "icosphere4":
        meshes, ratio, attrs = [W.icosphere(4)], (1, 10), False
    elif name == "seamed64":
        meshes, ratio, attrs = [W.seamed_grid(64, seams=4)], (1, 5), True
    else:
        meshes, ratio, attrs = W.patches(count=24, n=256), (1, 2), False
    for m in meshes:
        t = W.target_count(m.index_count, *ratio)
        A = m.vertices[:, 3:8] if attrs else None
        w = W.SEAMED_GRID_WEIGHTS if attrs else None
        opts = 1 if name == "patches" else 0
        r1, e1 = O.simplify(m.indices, m.vertices, t, 1e-2, opts, attributes=A, weights=w, impl="ref")
        r2, e2 = O.simplify(m.indices, m.vertices, t, 1e-2, opts, attributes=A, weights=w, rounds=True, window=1000)
        assert r1.tolist() == r2.tolist() and same_bits(e1, e2)


def test_oracle_trace_matches_reference_trace(capfd):
    """Per-pass counters of the oracle equal the reference's TRACE=1 lines
    (src/simplifier.cpp:1629-1632) on the 64x64 grid."""
    _ref_or_skip()
    import os
    if not os.path.exists(O.REF_TRACE_SO):
        pytest.skip("trace build missing")
    m = W.grid(64)
    t = W.target_count(m.index_count, 1, 4)
    import ctypes
    O.simplify(m.indices, m.vertices, t, 1e-2, 0, impl="ref_trace")
    ctypes.CDLL(None).fflush(None)
    out = capfd.readouterr().out
    lines = [l for l in out.splitlines() if "evaluated" in l]
    _, _, st = O.simplify(m.indices, m.vertices, t, 1e-2, 0, stats=True, rounds=True)
    passes = st.pass_list()
    assert len(lines) == len(passes)
    for line, p in zip(lines, passes):
        ev = line.split("evaluated ")[1].split(" collapses")[0]
        evaluated, cand = (int(x) for x in ev.split("/"))
        done = int(line.split("(done ")[1].split(",")[0])
        skipped = int(line.split("skipped ")[1].split(",")[0])
        invalid = int(line.split("invalid ")[1].split(")")[0])
        assert (evaluated, cand, done, skipped, invalid) == (p["evaluated"], p["candidates"], p["committed"], p["skipped"], p["rejected"])


def test_closed_forms():
    """Quantities fixed by the mathematics, independent of any implementation detail."""
    # simplifyScale is the largest bounding-box side (demo/tests.cpp:1943-1948)
    assert O.simplify_scale(np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 3]], np.float32)) == 3.0
    # a flat grid collapses with zero error down to the border-locked minimum
    m = W.grid(9)
    v = m.vertices.copy()
    v[:, 2] = 0
    r, e = O.simplify(m.indices, v, 0, 1e-2, 1)
    assert e < 1e-3
    # invariants: output references input vertices only; no triangle repeats a position
    rng = np.random.default_rng(7)
    ib, vv = random_soup(rng)
    r, e = O.simplify(ib, vv, 30, 0.5)
    assert r.size <= ib.size and r.size % 3 == 0
    assert set(r.tolist()) <= set(ib.tolist())


def test_plausible_mistakes_are_caught():
    """A planted mistake in the restatement's inputs must change the result: a transposed
    triangle winding flips every face normal, which changes the flip test outcome."""
    m = W.grid(16)
    t = W.target_count(m.index_count, 1, 4)
    r1, e1 = O.simplify(m.indices, m.vertices, t, 1e-2)
    v2 = m.vertices.copy()
    v2[:, 2] *= np.float32(1.0001)
    r2, e2 = O.simplify(m.indices, v2, t, 1e-2)
    assert not (r1.tolist() == r2.tolist() and same_bits(e1, e2))


SLOPPY_CASES = kat.load_cases("sloppy_known_answers.json")


@pytest.mark.parametrize("case", SLOPPY_CASES, ids=[c["name"] for c in SLOPPY_CASES])
def test_sloppy_known_answers(case):
    """meshopt_simplifySloppy known answers of demo/tests.cpp:1833-1882 on the C restatement."""
    a = kat.arrays(case)
    r, e = O.simplify_sloppy(a["indices"], a["positions"], case["target_index_count"], case["target_error"], lock=a["lock"])
    kat.check(case, r, e)


@pytest.mark.parametrize("seed", range(60))
def test_sloppy_oracle_matches_reference(seed):
    """The sloppy restatement against the compiled reference: index buffer and error bits."""
    _ref_or_skip()
    rng = np.random.default_rng(seed)
    if seed % 2:
        ib, v = random_soup(rng, nv=int(rng.integers(3, 200)), ntri=int(rng.integers(1, 400)))
    else:
        ib, v = random_grid(rng, n=int(rng.integers(3, 30)))
    v = np.ascontiguousarray(v[:, :3])
    target = (ib.size // 3 * int(rng.integers(0, 5)) // 4) * 3
    err = float(rng.choice([1e-4, 1e-3, 1e-2, 0.1, 1.0]))
    lock = (rng.random(v.shape[0]) < 0.1).astype(np.uint8) if seed % 3 == 0 else None
    r1, e1 = O.simplify_sloppy(ib, v, target, err, lock=lock)
    r2, e2 = O.simplify_sloppy(ib, v, target, err, lock=lock, impl="ref")
    assert np.array_equal(r1, r2)
    assert same_bits(e1, e2)


def test_sloppy_closed_form():
    """A flat (n+1)^2 grid with every triangle as the target and error 1e-3: the minimum grid is
    int(1.f / 1e-3f) = 999 cells per axis (1e-3f is a little above 0.001, so the quotient
    truncates to 999), where each vertex has a cell of its own, so the search
    stops before its first pass (the count already meets the target), there are V cells, and
    every triangle survives (rotated to lead with its smallest index)."""
    n = 16
    m = W.grid(n + 1)
    v = np.ascontiguousarray(m.vertices[:, :3]).copy()
    v[:, 2] = 0.0
    r, e, tr = O.simplify_sloppy(m.indices, v, m.index_count, 1e-3, trace=True)
    assert tr.passes == 0 and tr.final_grid == 999
    assert tr.cells == v.shape[0]
    assert r.size == m.index_count
    tri = m.indices.reshape(-1, 3)
    rot = [tuple(np.roll(t, -int(np.argmin(t)))) for t in tri]
    assert [tuple(x) for x in r.reshape(-1, 3)] == rot


UPDATE_CASES = kat.load_cases("update_known_answers.json")


def _update_run(ib, vb, na, target, err, opts, weights, lock, impl, rounds=False):
    """meshopt_simplifyWithUpdate on a copy of the vertex buffer vb (xyz, then na attribute
    columns); returns (indices, error, updated copy)."""
    b = vb.copy()
    r, e = O.simplify_with_update(ib, b[:, :3], target, err, opts, attributes=b[:, 3:3 + na] if na else None,
                                  weights=weights, lock=lock, impl=impl, solve_rounds=rounds)
    return r, e, b


@pytest.mark.parametrize("impl", ["oracle", "rounds", "ref"])
@pytest.mark.parametrize("case", UPDATE_CASES, ids=[c["name"] for c in UPDATE_CASES])
def test_update_known_answers(case, impl):
    """meshopt_simplifyWithUpdate known answers (demo/tests.cpp:2590-2670,
    js/meshopt_simplifier.test.js:115-138): index buffer, moved centre vertex, border vertices
    without drift, locked vertex untouched; in index order, in dependency rounds, and on the
    compiled reference."""
    if impl == "ref":
        _ref_or_skip()
    a = kat.arrays(case)
    r, e = O.simplify_with_update(a["indices"], a["positions"], case["target_index_count"], case["target_error"],
                                  case["options"], attributes=a["attributes"], weights=a["weights"], lock=a["lock"],
                                  impl="ref" if impl == "ref" else "oracle", solve_rounds=impl == "rounds")
    kat.check(case, r, e)
    kat.check_update(case, a["positions"], a["attributes"])


@pytest.mark.parametrize("seed", range(40))
def test_update_oracle_matches_reference(seed):
    """The solve restatement against the compiled reference on random meshes with every option
    set, 0-3 attributes (one of weight 0) and random locks: index buffer, error bits and every
    byte of the updated vertex buffer; the round schedule must agree bit for bit as well."""
    _ref_or_skip()
    args = update_case(seed)
    r1, e1, b1 = _update_run(*args, impl="ref")
    r2, e2, b2 = _update_run(*args, impl="oracle")
    r3, e3, b3 = _update_run(*args, impl="oracle", rounds=True)
    assert r1.tolist() == r2.tolist() == r3.tolist()
    assert same_bits(e1, e2) and same_bits(e1, e3)
    assert b1.tobytes() == b2.tobytes() == b3.tobytes()


@pytest.mark.parametrize("name,opts", [("grid64", 0), ("seamed64", 0), ("seamed64", 2), ("icosphere4", 8)])
def test_update_oracle_matches_reference_workloads(name, opts):
    """Workload meshes through the solve: a row-major grid (a long dependency chain), a seamed
    grid with two attribute columns (wedges, complex vertices, the volume constraint)."""
    _ref_or_skip()
    m = {"grid64": lambda: W.grid(64), "seamed64": lambda: W.seamed_grid(64, seams=4, seed=3),
         "icosphere4": lambda: W.icosphere(4)}[name]()
    vb = np.ascontiguousarray(m.vertices, dtype=np.float32)
    na = vb.shape[1] - 3
    w = np.ones(na, np.float32) if na else None
    args = (m.indices, vb, na, W.target_count(m.index_count, 1, 10), 1e-2, opts, w, None)
    r1, e1, b1 = _update_run(*args, impl="ref")
    r2, e2, b2 = _update_run(*args, impl="oracle", rounds=True)
    assert r1.tolist() == r2.tolist() and same_bits(e1, e2)
    assert b1.tobytes() == b2.tobytes()
    assert (b1[:, :3] != vb[:, :3]).any()  # the solve moved something


@pytest.mark.parametrize("seed", range(20))
def test_update_rounds_match_index_order(seed):
    """The device's schedule (dependency rounds, each solved from the positions at its start)
    replays the reference's index-order walk exactly; checked on the restatement alone."""
    args = update_case(1000 + seed)
    r1, e1, b1 = _update_run(*args, impl="oracle")
    r2, e2, b2 = _update_run(*args, impl="oracle", rounds=True)
    assert r1.tolist() == r2.tolist() and b1.tobytes() == b2.tobytes()


def test_update_closed_form():
    """Mathematics the solve must respect without any collapse (target = every index):
    every face plane of a vertex's fan passes through the vertex, so its quadric error is zero
    there and the optimum (with the point regulariser) is the vertex itself up to rounding; and
    a mesh whose vertices all lie on one line has zero-area faces, an all-zero quadric, a
    singular system, and keeps every vertex exactly (up to the rescale round trip)."""
    m = W.grid(9)
    vb = np.ascontiguousarray(m.vertices[:, :3], dtype=np.float32)
    r, e, b = _update_run(m.indices, vb, 0, m.index_count, 1e-2, 0, None, None, impl="oracle")
    assert r.size == m.index_count and e == 0
    assert np.allclose(b, vb, atol=1e-5)
    flat = vb.copy()
    flat[:, 1] = flat[:, 0]  # every vertex on the line x = y = z: zero-area faces
    flat[:, 2] = flat[:, 0]
    r, e, b = _update_run(m.indices, flat, 0, m.index_count, 1e-2, 0, None, None, impl="oracle")
    assert np.allclose(b, flat, atol=1e-5)


@pytest.mark.parametrize("seed", range(8))
def test_boundary_locks_rule(seed):
    """lockBoundary (demo/clusterlod.h:372-419) marks, group by group, every remapped vertex already
    seen by an earlier group (bit 7 as 'seen', bit 0 as the lock), then rewrites every vertex in
    index order; the device computes the same bytes as 'the position is used by two or more groups'
    (the lowest group that uses it and any other), plus the caller's Lock bit of the position's
    first vertex, plus the vertex's own caller bits.  Both forms on random group sets, and the form
    without the first vertex's Lock bit must differ from the loop somewhere."""
    from meshes import lock_boundary
    rng = np.random.default_rng(seed)
    V = int(rng.integers(20, 200))
    remap = np.arange(V)
    dup = rng.random(V) < 0.3
    remap[dup] = rng.integers(0, V, size=int(dup.sum()))
    remap = np.minimum(remap, np.arange(V))  # a remap points at a lower (or the same) index
    remap = remap[remap]                     # and at a primary
    groups = [rng.integers(0, V, size=3 * int(rng.integers(1, 30))) for _ in range(int(rng.integers(1, 12)))]
    base = rng.integers(0, 8, size=V).astype(np.uint8)
    prior = (rng.integers(0, 256, size=V).astype(np.uint8) & 2)  # earlier levels' protect bits

    ref = lock_boundary(remap, groups, base, prior)

    uses = np.zeros(V, np.int32)
    for grp in groups:
        uses[np.unique(remap[grp])] += 1
    boundary = (uses >= 2)[remap].astype(np.uint8)
    dev = boundary | (base[remap] & 1) | (prior & 2) | base
    assert np.array_equal(ref, dev)
    without_first = boundary | (prior & 2) | base
    gains = (remap != np.arange(V)) & ((base[remap] & 1) == 1) & ((base & 1) == 0) & (boundary == 0)
    assert np.array_equal(np.flatnonzero(ref != without_first), np.flatnonzero(gains))


@pytest.mark.parametrize("seed", range(40))
def test_oracle_matches_reference_edge_values(seed):
    """NaN, +-0 and denormal coordinates and attributes (tests/meshes.py edge_value_case): the
    restatement, its round-based commit and the reference library give the same bits, with and
    without the attribute column, across option sets."""
    _ref_or_skip()
    from meshes import edge_value_case
    ib, v, kind = edge_value_case(seed)
    rng = np.random.default_rng(seed)
    opts = OPTION_SETS[seed % len(OPTION_SETS)]
    target = (ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3
    err = float(rng.choice([1e-3, 1e-2, 1e-1, 1.0]))
    attrs = weights = None
    if seed % 3 == 0:
        attrs, weights = v[:, 3:4], np.array([0.5], np.float32)
    r1, e1 = O.simplify(ib, v, target, err, opts, attributes=attrs, weights=weights, impl="ref")
    r2, e2 = O.simplify(ib, v, target, err, opts, attributes=attrs, weights=weights)
    r3, e3 = O.simplify(ib, v, target, err, opts, attributes=attrs, weights=weights, rounds=True, window=5)
    assert r1.tolist() == r2.tolist() and same_bits(e1, e2), kind
    assert r1.tolist() == r3.tolist() and same_bits(e1, e3), kind
    s1, f1 = O.simplify_sloppy(ib, v, target, err, impl="ref")
    s2, f2 = O.simplify_sloppy(ib, v, target, err)
    assert s1.tolist() == s2.tolist() and same_bits(f1, f2), kind
    assert O.simplify_scale(v, impl="ref") == O.simplify_scale(v) or (np.isnan(O.simplify_scale(v)) and np.isnan(O.simplify_scale(v, impl="ref")))


@pytest.mark.parametrize("seed", range(20))
def test_update_edge_values(seed):
    """meshopt_simplifyWithUpdate on NaN, +-0 and denormal inputs: the restatement (index order and
    rounds) against the reference library, every byte of the updated vertex buffer included."""
    _ref_or_skip()
    from meshes import edge_value_case
    ib, v, kind = edge_value_case(seed)
    rng = np.random.default_rng(500 + seed)
    na = int(seed % 2)
    target = (ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3
    err = float(rng.choice([1e-2, 1e-1, 1.0]))
    opts = [0, 2, 8, 1][seed % 4]
    w = np.array([0.5], np.float32) if na else None
    args = (ib, v, na, target, err, opts, w, None)
    r1, e1, b1 = _update_run(*args, impl="ref")
    r2, e2, b2 = _update_run(*args, impl="oracle")
    r3, e3, b3 = _update_run(*args, impl="oracle", rounds=True)
    assert r1.tolist() == r2.tolist() == r3.tolist(), kind
    assert same_bits(e1, e2) and same_bits(e1, e3), kind
    assert b1.tobytes() == b2.tobytes() == b3.tobytes(), kind


@pytest.mark.slow
@pytest.mark.parametrize("name", ["grid256", "icosphere8", "seamed1024", "patches4096"])
def test_oracle_matches_reference_full_size(name):
    """The restatement against the reference library at the benchmark workloads' full sizes: the
    256x256 grid (target 25%), icosphere L8 (10%), the seamed 1024x1024 grid with normal and UV
    attributes (20%), and all 4096 patches of the batched workload (LockBorder, 50%); index
    buffers, counts and result_error bits.  The GPU's full-size parity tests compare against the
    restatement, so this is what ties them to the reference."""
    _ref_or_skip()
    if name == "patches4096":
        ms = W.patches(4096)
        meshes = [(m.indices, m.vertices) for m in ms]
        import meshoptimizer_amd as M
        idx, io, verts, vo, _ = M.pack_meshes(meshes)
        tg = np.array([W.target_count(m.index_count, 1, 2) for m in ms], np.uint32)
        d1, c1, e1 = O.simplify_batch(idx, io, verts, vo, tg, 1e-2, 1, threads=8, impl="ref")
        d2, c2, e2 = O.simplify_batch(idx, io, verts, vo, tg, 1e-2, 1, threads=8, impl="oracle")
        assert np.array_equal(c1, c2)
        assert e1.tobytes() == e2.tobytes()
        for p in range(len(ms)):
            assert np.array_equal(d1[io[p]:io[p] + c1[p]], d2[io[p]:io[p] + c2[p]]), f"patch {p}"
        assert int(c1.sum()) < int(io[-1])
        return
    if name == "grid256":
        m, ratio, attrs = W.grid(256), (1, 4), False
    elif name == "icosphere8":
        m, ratio, attrs = W.icosphere(8), (1, 10), False
    else:
        m, ratio, attrs = W.seamed_grid(1024), (1, 5), True
    t = W.target_count(m.index_count, *ratio)
    A = m.vertices[:, 3:8] if attrs else None
    w = W.SEAMED_GRID_WEIGHTS if attrs else None
    r1, e1 = O.simplify(m.indices, m.vertices, t, 1e-2, 0, attributes=A, weights=w, impl="ref")
    r2, e2 = O.simplify(m.indices, m.vertices, t, 1e-2, 0, attributes=A, weights=w)
    assert r1.size == r2.size and np.array_equal(r1, r2)
    assert same_bits(e1, e2)
    assert r1.size <= t or e1 > 0


def test_clod_edge_limit_closed_forms():
    """The clusterlod edge-length limit (demo/clusterlod.h:489-516) on triangles whose answer is
    fixed by hand: an equilateral triangle of side s gives s^2 (its shortest edge), a collinear
    wire of sides 1, 9 and 10 gives 10^2 / 4 (a quarter of its longest), the largest triangle
    wins in either order, a point triangle gives 0, and a triangle with a NaN vertex never raises
    the maximum."""
    from oracle import clod as C
    h = np.float32(np.sqrt(np.float32(3)) / 2)
    eq = np.array([[0, 0, 0], [2, 0, 0], [1, 2 * h, 0]], np.float32)
    assert C.edge_limit_sq([0, 1, 2], eq) == np.float32(4)
    wire = np.array([[0, 0, 0], [1, 0, 0], [10, 0, 0]], np.float32)
    assert C.edge_limit_sq([0, 1, 2], wire) == np.float32(25)
    both = np.concatenate([eq, wire])
    assert C.edge_limit_sq([0, 1, 2, 3, 4, 5], both) == np.float32(25)
    assert C.edge_limit_sq([3, 4, 5, 0, 1, 2], both) == np.float32(25)
    flat = np.array([[1, 1, 1], [1, 1, 1], [1, 1, 1]], np.float32)
    assert C.edge_limit_sq([0, 1, 2], flat) == np.float32(0)
    nan = np.concatenate([eq, np.array([[np.nan, 0, 0], [5, 0, 0], [0, 5, 0]], np.float32)])
    assert C.edge_limit_sq([3, 4, 5, 0, 1, 2], nan) == np.float32(4)
    assert C.edge_limit_sq([3, 4, 5], nan) == np.float32(0)


def test_clod_simplify_paths():
    """clusterlod's per-group simplify: a target above the group's size returns the group with
    error 0; with the sloppy fallback a heavily locked group that the edge collapser leaves above
    its target comes back at most at the target (the sloppy call's own bound) and its error is the
    sloppy error times the subset's extent times the factor; the edge limit caps the error at
    sqrt(edge_limit_sq) * limit."""
    from oracle import clod as C
    rng = np.random.default_rng(3)
    ib, v = random_grid(rng, n=12)
    locks = (rng.random(v.shape[0]) < 0.6).astype(np.uint8)
    r, e = C.clod_simplify(ib, v, locks, ib.size + 3, 4)
    assert np.array_equal(r, ib) and e == 0
    target = ib.size // 3 // 8 * 3
    r0, _ = C.clod_simplify(ib, v, locks, target, 4)
    assert r0.size > target
    cfg = dict(C.DEFAULT_CONFIG)
    r1, e1 = C.clod_simplify(ib, v, locks, target, 4, cfg)
    sub = np.ascontiguousarray(v[ib, :3])
    rs, es = O.simplify_sloppy(np.arange(ib.size, dtype=np.uint32), sub, target, C.FLT_MAX, lock=locks[ib])
    assert np.array_equal(r1, ib[rs])
    assert e1 == np.float32(np.float32(es) * np.float32(O.simplify_scale(sub))) * np.float32(2)
    cfg["error_edge_limit"] = 1e-3
    _, e2 = C.clod_simplify(ib, v, locks, target, 4, cfg)
    assert e2 == min(e1, np.float32(np.sqrt(C.edge_limit_sq(ib, v)) * np.float32(1e-3)))
    assert e2 < e1
