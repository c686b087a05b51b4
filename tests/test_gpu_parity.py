"""GPU parity: the HIP library against the oracle (and the reference's known answers),
0 ulp on the output index buffer, the index count and result_error."""
import sys

import numpy as np
import pytest

import kat
from meshes import edge_value_case, scattered_components, random_grid, random_soup, same_bits
from oracle import oracle as O
import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = kat.load_cases()


def announce(what):
    print(f"[case] {what}", flush=True)
    sys.stdout.flush()


def gpu_call(ib, v, target, err, opts, attrs=None, weights=None, lock=None, with_attributes=False):
    if with_attributes or attrs is not None or lock is not None or weights is not None:
        return M.simplify_with_attributes(ib, v, attrs, weights, lock, target, err, opts)
    return M.simplify(ib, v, target, err, opts)


def assert_parity(ib, v, target, err, opts, attrs=None, weights=None, lock=None, with_attributes=False):
    rg, eg = gpu_call(ib, v, target, err, opts, attrs, weights, lock, with_attributes)
    ro, eo = O.simplify(ib, v, target, err, opts, attributes=attrs, weights=weights, lock=lock)
    assert rg.size == ro.size, f"index count {rg.size} != oracle {ro.size}"
    assert np.array_equal(rg, ro), f"first difference at {int(np.argmax(rg != ro))}"
    assert same_bits(eg, eo), f"result_error {eg!r} != oracle {eo!r}"
    return rg, eg


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_known_answers(case):
    announce(case["name"])
    a = kat.arrays(case)
    with_attr = case["call"] == "simplifyWithAttributes"
    attrs = a["attributes"] if with_attr else None
    r, e = assert_parity(a["indices"], a["positions"], case["target_index_count"], case["target_error"], case["options"],
                         attrs, a["weights"], a["lock"], with_attributes=with_attr)
    kat.check(case, r, e)


OPTION_SETS = [0, 1, 2, 4, 16, 32, 64, 1 | 32, 32 | 16, 1 | 2 | 4, 8, 8 | 1, 8 | 32, 8 | 2 | 4]


@pytest.mark.parametrize("seed", range(60))
def test_random_meshes(seed):
    rng = np.random.default_rng(1000 + seed)
    opts = OPTION_SETS[seed % len(OPTION_SETS)]
    if seed % 2:
        ib, v = random_soup(rng, nv=int(rng.integers(20, 150)), ntri=int(rng.integers(10, 400)))
    else:
        ib, v = random_grid(rng, n=int(rng.integers(5, 24)))
    target = (ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3
    err = float(rng.choice([1e-3, 1e-2, 1e-1, 1.0]))
    attrs = weights = lock = None
    if seed % 3 == 0:
        attrs, weights = v[:, 3:4], np.array([0.5], np.float32)
    if seed % 5 == 0:
        lock = rng.integers(0, 8, size=v.shape[0]).astype(np.uint8)
        attrs, weights = v[:, 3:4], np.array([0.25], np.float32)
    announce(f"random seed={seed} opts={opts} I={ib.size} V={v.shape[0]}")
    assert_parity(ib, v, target, err, opts, attrs, weights, lock)


def test_destination_aliasing():
    """destination may alias indices (src/simplifier.cpp:2361-2363, demo/tests.cpp:2271)."""
    import ctypes
    case = next(c for c in CASES if c["name"] == "simplifySparse.geometry")
    a = kat.arrays(case)
    ib = a["indices"].copy()
    pos = a["positions"]
    lib = M.library()
    n = lib.meshopt_simplify(ib.ctypes.data, ib.ctypes.data, ib.size, pos.ctypes.data, pos.shape[0], pos.strides[0],
                             case["target_index_count"], case["target_error"], case["options"], None)
    assert ib[:n].tolist() == case["expected_indices"]


@pytest.mark.parametrize("name", ["grid64", "icosphere4", "seamed128", "patches"])
def test_workloads_small(name):
    if name == "grid64":
        meshes, ratio, attrs, opts = [W.grid(64)], (1, 4), False, 0
    elif name == "icosphere4":
        meshes, ratio, attrs, opts = [W.icosphere(4)], (1, 10), False, 0
    elif name == "seamed128":
        meshes, ratio, attrs, opts = [W.seamed_grid(128, seams=6)], (1, 5), True, 0
    else:
        meshes, ratio, attrs, opts = W.patches(count=16, n=256), (1, 2), False, M.SimplifyLockBorder
    for m in meshes:
        announce(f"{name} {m.name} T={m.triangle_count}")
        t = W.target_count(m.index_count, *ratio)
        A = m.vertices[:, 3:8] if attrs else None
        w = W.SEAMED_GRID_WEIGHTS if attrs else None
        assert_parity(m.indices, m.vertices, t, 1e-2, opts, A, w)


@pytest.mark.slow
@pytest.mark.parametrize("name", ["grid256", "icosphere8", "seamed1024"])
def test_workloads_full(name):
    if name == "grid256":
        m, ratio, attrs = W.grid(256), (1, 4), False
    elif name == "icosphere8":
        m, ratio, attrs = W.icosphere(8), (1, 10), False
    else:
        m, ratio, attrs = W.seamed_grid(1024), (1, 5), True
    announce(f"{name} T={m.triangle_count}")
    t = W.target_count(m.index_count, *ratio)
    A = m.vertices[:, 3:8] if attrs else None
    w = W.SEAMED_GRID_WEIGHTS if attrs else None
    assert_parity(m.indices, m.vertices, t, 1e-2, 0, A, w)
    # once the process has seen the call's window sizes: one read-back per pass, plus the index
    # check, the set-up and the result copy (a window size seen for the first time adds a read-back
    # per batch of its head rounds)
    assert_parity(m.indices, m.vertices, t, 1e-2, 0, A, w)
    call, passes = M.last_call_stats()
    assert call["read_backs"] <= len(passes) + 3, (call["read_backs"], len(passes))


@pytest.mark.parametrize("seed", range(8))
def test_prune_components(seed):
    """Prune (src/simplifier.cpp:1866-2024, 2457-2472, 2539-2569) on meshes of many components."""
    rng = np.random.default_rng(500 + seed)
    ib, v = scattered_components(rng, int(rng.integers(4, 40)))
    target = 0 if seed % 2 else (ib.size // 3 * int(rng.integers(1, 3)) // 4) * 3
    err = float(rng.choice([1e-3, 1e-2, 5e-2, 0.2]))
    opts = M.SimplifyPrune | int(rng.choice([0, M.SimplifyLockBorder, M.SimplifySparse, M.SimplifyErrorAbsolute]))
    announce(f"prune seed={seed} I={ib.size} V={v.shape[0]} opts={opts} err={err}")
    assert_parity(ib, v, target, err, opts)


@pytest.mark.parametrize("seed", range(40))
def test_edge_values(seed):
    """NaN, +-0 and denormal coordinates and attributes (tests/meshes.py edge_value_case): the
    position remap's -0 hashing and == welding (src/simplifier.cpp:111-148, 201-240), the box's
    first-of-+-0 and NaN-skipping compares (572-573), denormals kept through every kernel, and a
    denormal extent whose reciprocal overflows.  Same parameters as the oracle's pin against the
    reference library."""
    ib, v, kind = edge_value_case(seed)
    rng = np.random.default_rng(seed)
    opts = [0, 1, 2, 4, 16, 32, 64, 1 | 32, 2 | 8, 8, 1 | 8, 32 | 16][seed % 12]
    target = (ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3
    err = float(rng.choice([1e-3, 1e-2, 1e-1, 1.0]))
    attrs = weights = None
    if seed % 3 == 0:
        attrs, weights = v[:, 3:4], np.array([0.5], np.float32)
    announce(f"edge values seed={seed} kind={kind} opts={opts} I={ib.size} V={v.shape[0]}")
    assert_parity(ib, v, target, err, opts, attrs, weights)


def test_concurrent_calls():
    """meshopt_simplify from eight threads at once (each call takes its own context, stream and
    arena): every result equals the oracle's, as calls of the reference from several threads do."""
    import concurrent.futures as cf
    rng = np.random.default_rng(808)
    cases = []
    for k in range(24):
        ib, v = random_grid(rng, n=int(rng.integers(8, 40))) if k % 2 else random_soup(rng, nv=int(rng.integers(50, 400)), ntri=int(rng.integers(50, 800)))
        cases.append((ib, v, (ib.size // 3 * int(rng.integers(1, 4)) // 4) * 3, float(rng.choice([1e-2, 1e-1]))))
    announce(f"concurrent calls: {len(cases)} meshes on 8 threads")
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        got = list(ex.map(lambda c: M.simplify(c[0], c[1], c[2], c[3], 0), cases))
    for k, ((ib, v, t, e), (rg, eg)) in enumerate(zip(cases, got)):
        ro, eo = O.simplify(ib, v, t, e, 0)
        assert np.array_equal(rg, ro) and same_bits(eg, eo), f"case {k}: differs from the oracle"


@pytest.mark.parametrize("policy", [(-1, 0), (-1, 2), (1 << 20, 0)], ids=["deep-all", "deep-from-2", "cheap-only"])
def test_round_policies(policy):
    """Every commit-round policy gives the same result: deep rounds everywhere (the moving corners'
    alternatives examined from the first round and in every tail), deep from the third head round,
    and cheap rounds only; on random meshes of every option set (several windows of rounds for the
    larger ones), a seamed grid with attributes, and the full-size grid 256 workload (its first
    pass needs more head rounds than the first guess)."""
    rng = np.random.default_rng(77)
    M.set_round_policy(*policy)
    try:
        for k in range(14):
            if k % 2:
                ib, v = random_soup(rng, nv=int(rng.integers(20, 2000)), ntri=int(rng.integers(10, 4000)))
            else:
                ib, v = random_grid(rng, n=int(rng.integers(8, 64)))
            target = (ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3
            announce(f"round policy {policy} mesh {k} I={ib.size}")
            assert_parity(ib, v, target, float(rng.choice([1e-2, 1e-1, 1.0])), OPTION_SETS[k % len(OPTION_SETS)])
        m = W.seamed_grid(128, seams=6)
        announce(f"round policy {policy} seamed128 with attributes")
        assert_parity(m.indices, m.vertices, W.target_count(m.index_count, 1, 5), 1e-2, 0, m.vertices[:, 3:8], W.SEAMED_GRID_WEIGHTS)
        g = W.grid(256)
        announce(f"round policy {policy} grid256")
        for _ in range(2):
            assert_parity(g.indices, g.vertices, W.target_count(g.index_count, 1, 4), 1e-2, 0)
    finally:
        M.set_round_policy()


@pytest.mark.parametrize("den,err,opts", [(20, 1e-2, 0), (100, 1.0, M.SimplifyLockBorder)])
def test_window_extension(den, err, opts):
    """A pass whose first rank window holds no break goes on to the rest of the list on the device
    (the extension window, with its own head rounds and tail): icosphere L6 at a low target opens
    one; the result matches the oracle."""
    m = W.icosphere(6)
    target = W.target_count(m.index_count, 1, den)
    announce(f"window extension icosphere6 target 1/{den} err {err} opts {opts}")
    assert_parity(m.indices, m.vertices, target, err, opts)
    call, passes = M.last_call_stats()
    assert call["device_extensions"] >= 1, (call, len(passes))


@pytest.mark.slow
def test_cold_round_counts():
    """A process that has not seen a mesh's window sizes runs their head rounds in batches with a
    read-back of the list length; a long-chained mesh (a 998 x 998 row-major grid) goes on to deep
    rounds within its first window, with the per-source candidate lists built then.  The cold call
    and the warm one after it both match the oracle."""
    g = W.grid(998)
    target = W.target_count(g.index_count, 3, 4)
    M.reset_round_counts()
    announce("cold grid998 target 3/4")
    assert_parity(g.indices, g.vertices, target, 1e-2, 0)
    call, passes = M.last_call_stats()
    print("[cold]", call["read_backs"], [(p["rounds"], p["tail_rounds"]) for p in passes], flush=True)
    assert call["read_backs"] > len(passes) + 3  # the batches of head rounds read the list back
    announce("warm grid998 target 3/4")
    assert_parity(g.indices, g.vertices, target, 1e-2, 0)
    call, passes = M.last_call_stats()
    assert call["read_backs"] <= len(passes) + 3
