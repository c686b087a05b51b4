"""Writes tests/golden/known_answers.json: the simplifier known-answer vectors held by the
reference's own tests, as data.

Sources (inputs and expected outputs transcribed as values):
  demo/tests.cpp:1771-2558           C known-answer tests of meshopt_simplify* (1.2)
  demo/tests.cpp:1833-1882           meshopt_simplifySloppy (written to sloppy_known_answers.json)
  demo/tests.cpp:2590-2670,          meshopt_simplifyWithUpdate (written to update_known_answers.json)
  js/meshopt_simplifier.test.js:115-138
  js/meshopt_simplifier.test.js:25-163  the WASM build's plain-text expected arrays

Each case: call ("simplify" | "simplifyWithAttributes"), vertices [V][F] float32 values,
optional attribute source ("interleaved" column offset or a separate [V][A] array),
weights, lock, indices, target_index_count, target_error, options, expected_count,
expected_indices (null = count-only check), and an optional error check.

Run:  python tests/golden/make_known_answers.py
"""
from __future__ import annotations

import json
import os

import numpy as np

LOCK_BORDER = 1
SPARSE = 2
ERROR_ABSOLUTE = 4
PRUNE = 8
DEBUG = 1 << 30

f32 = np.float32


def case(name, source, vertices, indices, target, error, options=0, *, call="simplify", vertex_floats=3,
         attr=None, weights=None, lock=None, expected_count, expected=None, err_lt=None, err_approx=None):
    v = np.asarray(vertices, dtype=np.float32).reshape(-1, vertex_floats)
    out = {
        "name": name,
        "source": source,
        "call": call,
        "vertices": [[float(x) for x in row] for row in v],
        "indices": [int(i) for i in indices],
        "target_index_count": int(target),
        "target_error": float(f32(error)),
        "options": int(options),
        "expected_count": int(expected_count),
        "expected_indices": None if expected is None else [int(i) for i in expected],
    }
    if attr is not None:
        if isinstance(attr, dict):
            out["attributes"] = attr
        else:
            a = np.asarray(attr, dtype=np.float32)
            out["attributes"] = {"separate": [[float(x) for x in row] for row in a.reshape(v.shape[0], -1)]}
    if weights is not None:
        out["weights"] = [float(f32(w)) for w in weights]
    if lock is not None:
        out["lock"] = [int(x) for x in lock]
    if err_lt is not None:
        out["error_lt"] = err_lt
    if err_approx is not None:
        out["error_approx"] = err_approx
    return out


def attr_grid(zstep_odd, js):
    """8x3 colour-gradient grid of the simplifyAttr tests (demo/tests.cpp:2027-2082 in float32,
    js/meshopt_simplifier.test.js:72-116 in float64 rounded once)."""
    vb = np.zeros((24, 6), dtype=np.float32)
    for y in range(8):
        if js:
            r = 0.8 + y * 0.05 if y < 4 else 0.0
            b = 0.0 if y < 4 else 0.8 + (7 - y) * 0.05
        else:
            r = f32(0.8) + f32(y) * f32(0.05) if y < 4 else f32(0)
            b = f32(0) if y < 4 else f32(0.8) + f32(7 - y) * f32(0.05)
        for x in range(3):
            if js:
                z = 0.03 * x + zstep_odd * (y % 2) + (1 if (x == 2 and y == 7) else 0) * 0.03
            else:
                z = f32(0.03) * f32(x) + f32(zstep_odd) * f32(y % 2) + f32(1 if (x == 2 and y == 7) else 0) * f32(0.03)
            vb[y * 3 + x] = [x, y, z, r, r, b]
    ib = []
    for y in range(7):
        for x in range(2):
            ib += [(y + 0) * 3 + (x + 0), (y + 0) * 3 + (x + 1), (y + 1) * 3 + (x + 0),
                   (y + 1) * 3 + (x + 0), (y + 0) * 3 + (x + 1), (y + 1) * 3 + (x + 1)]
    return vb, ib


def build():
    cases = []

    fan_ib = [0, 2, 1, 1, 2, 3, 3, 2, 4, 2, 5, 4]
    fan_vb = [0, 4, 0, 0, 1, 0, 2, 2, 0, 0, 0, 0, 1, 0, 0, 4, 0, 0]
    cases.append(case("simplify", "demo/tests.cpp:1771-1802", fan_vb, fan_ib, 3, 1e-2,
                      expected_count=3, expected=[0, 5, 3], err_lt=1e-4))

    tet_vb = [0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1]
    tet_ib = [0, 1, 2, 0, 2, 3, 0, 3, 1, 2, 1, 3]
    cases.append(case("simplifyStuck.tetrahedron", "demo/tests.cpp:1806-1810", tet_vb, tet_ib, 6, 1e-3,
                      expected_count=12, expected=tet_ib))
    strip_vb = [0, 0, 0, 1, 0, 0, 2, 0, 0, 0.5, 1, 0, 1.5, 1, 0]
    cases.append(case("simplifyStuck.strip", "demo/tests.cpp:1812-1818", strip_vb, [0, 1, 3, 3, 1, 4, 1, 2, 4], 6, 1e-3,
                      expected_count=6))
    cases.append(case("simplifyStuck.flipped_strip", "demo/tests.cpp:1812-1818", strip_vb, [0, 1, 3, 1, 3, 4, 1, 2, 4], 6, 1e-3,
                      expected_count=9))
    cases.append(case("simplifyStuck.quad", "demo/tests.cpp:1820-1824", [0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0],
                      [0, 1, 3, 0, 3, 2], 3, 1e-3, expected_count=6))
    cases.append(case("simplifyStuck.quad_split", "demo/tests.cpp:1826-1830",
                      [0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 1, 1, 0], [0, 1, 4, 0, 3, 2], 3, 1e-3, expected_count=6))

    flip_vb = [1, 1, -1, 1, 1, 1, 1, -1, 1, 1, -0.2, -0.2, 1, 0.2, -0.2, 1, -0.2, 0.2, 1, 0.2, 0.2, 1, 0.5, -0.5, 1, -1, 0]
    flip_ib = [7, 4, 3, 1, 2, 5, 7, 1, 6, 7, 8, 0, 7, 6, 4, 8, 5, 2, 8, 7, 3, 8, 3, 5, 5, 6, 1, 7, 0, 1]
    flip_exp = [0, 4, 3, 1, 2, 5, 0, 1, 6, 0, 6, 4, 8, 5, 2, 8, 0, 3, 8, 3, 5, 5, 6, 1]
    cases.append(case("simplifyFlip", "demo/tests.cpp:1894-1941", flip_vb, flip_ib, 3, 1e-3,
                      expected_count=24, expected=flip_exp))

    deg_vb = [0, 0, 0, 0, 1, 0, 0, 2, 0, 1, 0, 0, 2, 0, 0, 1, 1, 0]
    deg_ib = [0, 1, 3, 3, 1, 5, 1, 2, 5, 3, 5, 4, 1, 0, 1, 0, 3, 0]
    cases.append(case("simplifyDegenerate", "demo/tests.cpp:1950-1983", deg_vb, deg_ib, 3, 1e-3,
                      expected_count=6, expected=[0, 1, 4, 4, 1, 2]))

    g3_vb = [0, 0, 0, 0, 1, 0, 0, 2, 0, 1, 0, 0, 1, 1, 0, 1, 2, 0, 2, 0, 0, 2, 1, 0, 2, 2, 0]
    g3_ib = [0, 1, 3, 3, 1, 4, 1, 2, 4, 4, 2, 5, 3, 4, 6, 6, 4, 7, 4, 5, 7, 7, 5, 8]
    g3_exp = [0, 1, 3, 1, 2, 3, 3, 2, 5, 6, 3, 7, 3, 5, 7, 7, 5, 8]
    cases.append(case("simplifyLockBorder", "demo/tests.cpp:1985-2025", g3_vb, g3_ib, 3, 1e-3, LOCK_BORDER,
                      expected_count=18, expected=g3_exp))

    avb, aib = attr_grid(0.03, js=False)
    attr_exp = [0, 2, 11, 0, 11, 9, 9, 11, 12, 12, 11, 14, 12, 14, 23, 12, 23, 21]
    for skip_g in (False, True):
        cases.append(case(f"simplifyAttr(skip_g={skip_g})", "demo/tests.cpp:2027-2082", avb, aib, 18, 1e-2,
                          call="simplifyWithAttributes", vertex_floats=6, attr={"interleaved": 3, "count": 3},
                          weights=[0.5, 0.0 if skip_g else 0.5, 0.5], expected_count=18, expected=attr_exp))

    lock9 = [1, 1, 1, 1, 0, 1, 1, 1, 1]
    cases.append(case("simplifyLockFlags", "demo/tests.cpp:2084-2130", g3_vb, g3_ib, 3, 1e-3,
                      call="simplifyWithAttributes", lock=lock9, expected_count=18, expected=g3_exp))

    seam12_vb = [0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 2, 0, 1, 0, 0, 1, 1, 0, 1, 1, 0, 1, 2, 0, 2, 0, 0, 2, 1, 0, 2, 1, 0, 2, 2, 0]
    seam12_ib = [0, 1, 4, 4, 1, 5, 4, 5, 8, 8, 5, 9, 2, 3, 6, 6, 3, 7, 6, 7, 10, 10, 7, 11]
    locks = {
        None: 0,
        "lock0": ([1, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1], 12),
        "lock1": ([1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1], 18),
        "lock2": ([1, 0, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1], 24),
        "lock3": ([1, 1, 0, 1, 1, 1, 0, 1, 1, 1, 0, 1], 24),
    }
    cases.append(case("simplifyLockFlagsSeam.none", "demo/tests.cpp:2132-2202", seam12_vb, seam12_ib, 0, 1.0,
                      call="simplifyWithAttributes", expected_count=0))
    for lname in ("lock0", "lock1", "lock2", "lock3"):
        lk, cnt = locks[lname]
        cases.append(case(f"simplifyLockFlagsSeam.{lname}", "demo/tests.cpp:2132-2202", seam12_vb, seam12_ib, 0, 1.0,
                          call="simplifyWithAttributes", lock=lk, expected_count=cnt))

    sp_vb = [0, 0, 100, 0, 1, 0, 0, 2, 100, 1, 0, 0.1, 1, 1, 0.1, 1, 2, 0.1, 2, 0, 100, 2, 1, 0, 2, 2, 100]
    sp_vba = [[100], [0.5], [100], [0.5], [0.5], [0], [100], [0.5], [100]]
    sp_lock = [8, 1, 8, 1, 0, 1, 8, 1, 8]
    sp_ib = [3, 1, 4, 1, 5, 4, 3, 4, 7, 4, 5, 7]
    cases.append(case("simplifySparse.geometry", "demo/tests.cpp:2204-2273", sp_vb, sp_ib, 6, 1e-3, SPARSE,
                      expected_count=6, expected=[1, 5, 3, 3, 5, 7]))
    cases.append(case("simplifySparse.attributes", "demo/tests.cpp:2204-2273", sp_vb, sp_ib, 6, 1e-1, SPARSE,
                      call="simplifyWithAttributes", attr=sp_vba, weights=[0.5], lock=sp_lock,
                      expected_count=6, expected=[3, 1, 7, 1, 5, 7]))

    ea_vb = [0, 0, 0, 0, 1, 0, 0, 2, 0, 1, 0, 0, 1, 1, 1, 1, 2, 0, 2, 0, 0, 2, 1, 0, 2, 2, 0]
    cases.append(case("simplifyErrorAbsolute", "demo/tests.cpp:2275-2307", ea_vb, g3_ib, 18, 2.0,
                      LOCK_BORDER | ERROR_ABSOLUTE, expected_count=18, err_approx=[0.85, 0.01]))

    seam_vb = [0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 0, 2, 0, 1,
               1, 0, 0, 0, 1, 1, 0.3, 0, 1, 1, 0.3, 1, 1, 2, 0, 1,
               2, 0, 0, 0, 2, 1, 0.1, 0, 2, 1, 0.1, 1, 2, 2, 0, 1,
               3, 0, 0, 0, 3, 1, 0, 0, 3, 1, 0, 1, 3, 2, 0, 1]
    seam_ib = [0, 1, 4, 4, 1, 5, 2, 3, 6, 6, 3, 7, 4, 5, 8, 8, 5, 9, 6, 7, 10, 10, 7, 11, 8, 9, 12, 12, 9, 13, 10, 11, 14, 14, 11, 15]
    seam_exp = [0, 1, 13, 2, 3, 14, 0, 13, 12, 14, 3, 15]
    cases.append(case("simplifySeam.geometry", "demo/tests.cpp:2309-2374", seam_vb, seam_ib, 12, 1.0, vertex_floats=4,
                      expected_count=12, expected=seam_exp, err_approx=[0.1, 0.01]))
    cases.append(case("simplifySeam.attributes", "demo/tests.cpp:2309-2374", seam_vb, seam_ib, 12, 2.0, vertex_floats=4,
                      call="simplifyWithAttributes", attr={"interleaved": 3, "count": 1}, weights=[1.0],
                      expected_count=12, expected=seam_exp, err_approx=[0.1, 0.01]))

    fake_vb = [0, 0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 2, 0, 0, 0, 3]
    cases.append(case("simplifySeamFake", "demo/tests.cpp:2376-2392", fake_vb, [0, 1, 2, 2, 1, 3], 0, 1.0, vertex_floats=4,
                      expected_count=6))

    sa_vb = [0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 2, 0, 0,
             1, 0, 0, 1, 1, 1, 0, 1, 1, 1, 0, 1, 1, 2, 0, 1,
             4, 0, 0, 2, 4, 1, 0, 2, 4, 1, 0, 2, 4, 2, 0, 2]
    sa_ib = [0, 1, 4, 4, 1, 5, 2, 3, 6, 6, 3, 7, 4, 5, 8, 8, 5, 9, 6, 7, 10, 10, 7, 11]
    sa_exp = [0, 1, 4, 2, 3, 7, 4, 1, 8, 8, 1, 9, 2, 7, 10, 10, 7, 11]
    cases.append(case("simplifySeamAttr", "demo/tests.cpp:2394-2447", sa_vb, sa_ib, 12, 2.0, LOCK_BORDER, vertex_floats=4,
                      call="simplifyWithAttributes", attr={"interleaved": 3, "count": 1}, weights=[1.0],
                      expected_count=18, expected=sa_exp, err_approx=[0.35, 0.01]))

    cases.append(case("simplifyDebug", "demo/tests.cpp:2449-2482", fan_vb, fan_ib, 3, 1e-2, DEBUG,
                      expected_count=3, expected=[0 | (9 << 28), 5 | (9 << 28), 3 | (9 << 28)], err_lt=1e-4))

    pr_vb = [0, 4, 0, 0, 1, 0, 2, 2, 0, 0, 0, 0, 1, 0, 0, 4, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1]
    pr_ib = [3, 2, 4, 0, 2, 1, 1, 2, 3, 2, 5, 4, 6, 7, 8]
    cases.append(case("simplifyPrune", "demo/tests.cpp:2484-2526", pr_vb, pr_ib, 3, 1e-2, PRUNE,
                      expected_count=3, expected=[0, 5, 3], err_lt=1e-4))
    cases.append(case("simplifyPrune.subset", "demo/tests.cpp:2484-2526", pr_vb, [0, 5, 3], 3, 1e-2, PRUNE,
                      expected_count=3, expected=[0, 5, 3]))
    cases.append(case("simplifyPrune.subset_sparse", "demo/tests.cpp:2484-2526", pr_vb, [0, 5, 3], 3, 1e-2, PRUNE | SPARSE,
                      expected_count=3, expected=[0, 5, 3]))

    pc_vb = [0, 0, 0, 0, 1, 0, 1, 0, 0, 0, 0, 1, 0, 2, 1, 2, 0, 1, 0, 0, 2, 0, 4, 2, 4, 0, 2]
    cases.append(case("simplifyPruneCleanup", "demo/tests.cpp:2528-2558", pc_vb, list(range(9)), 3, 1.0,
                      LOCK_BORDER | PRUNE, expected_count=3, expected=[6, 7, 8], err_approx=[0.37, 0.01]))

    # WASM build's expected arrays (js/meshopt_simplifier.test.js)
    cases.append(case("js.simplify", "js/meshopt_simplifier.test.js:25-40", fan_vb, fan_ib, 3, 0.01,
                      expected_count=3, expected=[0, 5, 3], err_lt=1e-4))
    lb_vb = [0, 2, 0, 0, 1, 0, 1, 1, 0, 0, 0, 0, 1, 0, 0, 2, 0, 0]
    cases.append(case("js.simplifyLockBorder", "js/meshopt_simplifier.test.js:58-72", lb_vb, fan_ib, 3, 0.01, LOCK_BORDER,
                      expected_count=12, expected=fan_ib, err_lt=1e-4))
    jvb, jib = attr_grid(0.028, js=True)
    cases.append(case("js.simplifyAttr", "js/meshopt_simplifier.test.js:74-116", jvb, jib, 18, 1e-2,
                      call="simplifyWithAttributes", vertex_floats=6, attr={"interleaved": 3, "count": 3},
                      weights=[0.5, 0.5, 0.5], expected_count=18, expected=attr_exp))
    cases.append(case("js.simplifyLockFlags", "js/meshopt_simplifier.test.js:141-155", lb_vb, fan_ib, 3, 0.01,
                      call="simplifyWithAttributes", lock=[1, 1, 1, 1, 0, 1],
                      expected_count=9, expected=[0, 2, 1, 1, 2, 3, 2, 5, 3], err_lt=1e-4))
    return cases


def build_sloppy():
    """meshopt_simplifySloppy known answers (demo/tests.cpp:1833-1882)."""
    cases = []
    zero_vb = [0, 0, 0, 0, 0, 0, 0, 0, 0]
    cases.append(case("simplifySloppyStuck.zero_target", "demo/tests.cpp:1833-1841", zero_vb, [0, 1, 2], 0, 0.0,
                      call="simplifySloppy", expected_count=0))
    cases.append(case("simplifySloppyStuck.degenerate", "demo/tests.cpp:1843-1844", zero_vb, [0, 1, 2, 0, 1, 2], 6, 0.0,
                      call="simplifySloppy", expected_count=0))
    fan_vb = [0, 4, 0, 0, 1, 0, 2, 2, 0, 0, 0, 0, 1, 0, 0, 4, 0, 0]
    fan_ib = [0, 2, 1, 1, 2, 3, 3, 2, 4, 2, 5, 4]
    c = case("simplifySloppyLocks", "demo/tests.cpp:1847-1882", fan_vb, fan_ib, 3, 1.0, call="simplifySloppy",
             lock=[1, 0, 1, 0, 0, 1], expected_count=6, expected=[0, 2, 1, 1, 2, 5])
    c["error_eq"] = 0.0
    cases.append(c)
    return cases


def build_update():
    """meshopt_simplifyWithUpdate known answers: demo/tests.cpp:2590-2670 (simplifyUpdate,
    simplifyUpdateLocked with options 0 and Sparse) and js/meshopt_simplifier.test.js:115-138 (the
    same mesh with separate position and attribute arrays).  "update_checks" hold the tests'
    assertions on the updated vertex buffer: rows of the vertex array (positions then the
    attribute column), each "approx" [value, tolerance], "abs_lt" tolerance or "eq" value."""
    vb = [[0, 0, 0, 0], [1, 1, 0, 0], [2, 0, 0, 0], [0.9, 0.2, 0.1, 0.2], [1.1, 0.2, 0.1, 0.1]]
    ib = [0, 1, 3, 3, 1, 4, 4, 1, 2, 0, 3, 2, 3, 4, 2]
    expected = [0, 1, 3, 3, 1, 2, 0, 3, 2]
    border = [{"row": i, "col": 3, "abs_lt": 1e-6} for i in range(3)]
    moved = [{"row": 3, "col": c, "approx": [v, 1e-2]} for c, v in enumerate((0.88, 0.19, 0.11, 0.18))]
    exact = [{"row": 3, "col": c, "eq": float(f32(v))} for c, v in enumerate((0.9, 0.2, 0.1, 0.2))]
    cases = []
    c = case("simplifyUpdate", "demo/tests.cpp:2590-2626", vb, ib, 9, 1.0, call="simplifyWithUpdate", vertex_floats=4,
             attr={"interleaved": 3, "count": 1}, weights=[1.0], expected_count=9, expected=expected)
    c["update_checks"] = border + moved
    cases.append(c)
    for name, opts in (("simplifyUpdateLocked", 0), ("simplifyUpdateLocked.sparse", SPARSE)):
        c = case(name, "demo/tests.cpp:2628-2670, 3654-3655", vb, ib, 9, 1.0, opts, call="simplifyWithUpdate",
                 vertex_floats=4, attr={"interleaved": 3, "count": 1}, weights=[1.0], lock=[0, 0, 0, 1, 0],
                 expected_count=9, expected=expected)
        c["update_checks"] = border + exact
        cases.append(c)
    pos = [row[:3] for row in vb]
    c = case("js.simplifyUpdate", "js/meshopt_simplifier.test.js:115-138", pos, ib, 9, 1.0, call="simplifyWithUpdate",
             attr=[[row[3]] for row in vb], weights=[1.0], expected_count=9, expected=expected)
    # the separate attribute array is column 3 of the checked rows
    c["update_checks"] = border + moved
    cases.append(c)
    return cases


def main():
    cases = build()
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "known_answers.json")
    with open(path, "w") as f:
        json.dump({"reference": "meshoptimizer 1.2 (MESHOPTIMIZER_VERSION 1020)", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {path}")
    sloppy = build_sloppy()
    path = os.path.join(here, "sloppy_known_answers.json")
    with open(path, "w") as f:
        json.dump({"reference": "meshoptimizer 1.2 (MESHOPTIMIZER_VERSION 1020)", "cases": sloppy}, f, indent=1)
    print(f"wrote {len(sloppy)} cases to {path}")
    update = build_update()
    path = os.path.join(here, "update_known_answers.json")
    with open(path, "w") as f:
        json.dump({"reference": "meshoptimizer 1.2 (MESHOPTIMIZER_VERSION 1020)", "cases": update}, f, indent=1)
    print(f"wrote {len(update)} cases to {path}")


if __name__ == "__main__":
    main()
