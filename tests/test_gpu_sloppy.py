"""GPU parity of meshopt_simplifySloppy (src/simplifier.cpp:2647-2778): the HIP path against the
oracle (the C restatement, itself pinned to the compiled reference in test_oracle.py) on the
reference's known answers, random meshes and full-size workloads; index buffer, index count
and result_error bits must be identical."""
import sys

import numpy as np
import pytest

import kat
from meshes import edge_value_case, random_grid, random_soup, same_bits
from oracle import oracle as O
import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = kat.load_cases("sloppy_known_answers.json")


def announce(what):
    print(f"[case] {what}", flush=True)
    sys.stdout.flush()


def assert_parity(ib, v, target, err, lock=None):
    rg, eg = M.simplify_sloppy(ib, v, target, err, lock=lock)
    ro, eo = O.simplify_sloppy(ib, v, target, err, lock=lock)
    assert rg.size == ro.size, f"index count {rg.size} != oracle {ro.size}"
    assert np.array_equal(rg, ro), f"first difference at {int(np.argmax(rg != ro))}"
    assert same_bits(eg, eo), f"result_error {eg!r} != oracle {eo!r}"
    return rg, eg


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_sloppy_known_answers(case):
    announce(case["name"])
    a = kat.arrays(case)
    r, e = assert_parity(a["indices"], a["positions"], case["target_index_count"], case["target_error"], a["lock"])
    kat.check(case, r, e)


@pytest.mark.parametrize("seed", range(60))
def test_sloppy_random(seed):
    rng = np.random.default_rng(3000 + seed)
    if seed % 2:
        ib, v = random_soup(rng, nv=int(rng.integers(3, 300)), ntri=int(rng.integers(1, 600)))
    else:
        ib, v = random_grid(rng, n=int(rng.integers(3, 40)))
    v = np.ascontiguousarray(v[:, :3])
    target = (ib.size // 3 * int(rng.integers(0, 5)) // 4) * 3
    err = float(rng.choice([1e-4, 1e-3, 1e-2, 0.1, 1.0]))
    lock = (rng.random(v.shape[0]) < 0.1).astype(np.uint8) if seed % 3 == 0 else None
    announce(f"sloppy random seed={seed} I={ib.size} target={target} err={err} lock={lock is not None}")
    assert_parity(ib, v, target, err, lock)


@pytest.mark.parametrize("name,den,err", [("grid256", 4, 1e-2), ("grid256", 20, 1e-1), ("icosphere6", 10, 1e-2),
                                          ("icosphere6", 50, 0.2), ("seamed256", 5, 2e-2)])
def test_sloppy_workloads(name, den, err):
    m = {"grid256": lambda: W.grid(256), "icosphere6": lambda: W.icosphere(6),
         "seamed256": lambda: W.seamed_grid(256, seams=8, seed=5)}[name]()
    target = W.target_count(m.index_count, 1, den)
    announce(f"sloppy {name} target 1/{den} err {err}")
    v = np.ascontiguousarray(m.vertices[:, :3])
    r, e = assert_parity(m.indices, v, target, err)
    assert 0 < r.size <= m.index_count


def test_sloppy_device_entry():
    """meshopt_hip_simplify_sloppy_device on device-resident arrays gives the host entry's result,
    and rejects an out-of-range index with an error instead of reading past the positions."""
    from devmem import DeviceArray
    m = W.grid(64)
    v = np.ascontiguousarray(m.vertices[:, :3])
    target = W.target_count(m.index_count, 1, 4)
    announce("sloppy device entry")
    bad = m.indices.copy()
    bad[5] = v.shape[0] + 7
    with DeviceArray(m.indices) as d_ib, DeviceArray(v) as d_v, DeviceArray(nbytes=m.indices.nbytes) as d_dst, DeviceArray(bad) as d_bad:
        n, e = M.simplify_sloppy_device(d_dst.ptr, d_ib.ptr, m.index_count, d_v.ptr, v.shape[0], 12, None, target, 1e-2)
        rg = d_dst.to_host(np.uint32, n)
        rh, eh = M.simplify_sloppy(m.indices, v, target, 1e-2)
        assert np.array_equal(rg, rh) and same_bits(e, eh)
        with pytest.raises(RuntimeError, match="vertex_count"):
            M.simplify_sloppy_device(d_dst.ptr, d_bad.ptr, bad.size, d_v.ptr, v.shape[0], 12, None, target, 1e-2)


@pytest.mark.parametrize("seed", range(20))
def test_sloppy_edge_values(seed):
    """NaN, +-0 and denormal positions: cell ids of NaN and denormal coordinates, and the rule that
    a NaN error at a cell's first vertex keeps that vertex (the reference's comparisons)."""
    ib, v, kind = edge_value_case(seed)
    rng = np.random.default_rng(seed)
    v = np.ascontiguousarray(v[:, :3])
    target = (ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3
    err = float(rng.choice([1e-3, 1e-2, 1e-1, 1.0]))
    announce(f"sloppy edge values seed={seed} kind={kind} I={ib.size} target={target} err={err}")
    assert_parity(ib, v, target, err)
