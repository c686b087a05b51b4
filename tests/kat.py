"""Helpers that turn the known-answer fixtures (tests/golden/known_answers.json) into arrays."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_cases(name="known_answers.json"):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)["cases"]


def arrays(case):
    """Returns dict(indices, positions, attributes, weights, lock) ready for a simplify call.
    positions/attributes are float32 row views (row stride = vertex / attribute stride)."""
    v = np.asarray(case["vertices"], dtype=np.float32)
    positions = v
    attributes = None
    a = case.get("attributes")
    if a is not None:
        if "interleaved" in a:
            off = a["interleaved"]
            attributes = v[:, off:off + a["count"]]
        else:
            attributes = np.asarray(a["separate"], dtype=np.float32)
    weights = None if case.get("weights") is None else np.asarray(case["weights"], dtype=np.float32)
    lock = None if case.get("lock") is None else np.asarray(case["lock"], dtype=np.uint8)
    if case["call"] == "simplifyWithAttributes" and attributes is None:
        attributes = np.zeros((v.shape[0], 0), dtype=np.float32)
    return dict(indices=np.asarray(case["indices"], dtype=np.uint32), positions=positions,
                attributes=attributes, weights=weights, lock=lock)


def check(case, result, error):
    """Asserts the reference test's expectations on (result indices, result_error)."""
    name = case["name"]
    assert result.size == case["expected_count"], f"{name}: count {result.size} != {case['expected_count']}"
    if case["expected_indices"] is not None:
        assert result.tolist() == case["expected_indices"], f"{name}: {result.tolist()} != {case['expected_indices']}"
    if "error_eq" in case:
        assert error == case["error_eq"], f"{name}: error {error} != {case['error_eq']}"
    if "error_lt" in case:
        assert error < case["error_lt"], f"{name}: error {error}"
    if "error_approx" in case:
        val, tol = case["error_approx"]
        assert abs(error - val) < tol, f"{name}: error {error} vs {val}"


def updated_rows(positions, attributes):
    """The vertex rows an update check reads: xyz then the attribute columns."""
    pos = np.asarray(positions)[:, :3]
    if attributes is None or np.asarray(attributes).size == 0:
        return np.array(pos, dtype=np.float32)
    return np.concatenate([pos, np.asarray(attributes)], axis=1).astype(np.float32)


def check_update(case, positions, attributes):
    """Asserts the reference test's expectations on the vertex buffer a
    meshopt_simplifyWithUpdate call left behind."""
    rows = updated_rows(positions, attributes)
    for c in case.get("update_checks", []):
        v = float(rows[c["row"], c["col"]])
        where = f"{case['name']}: vertex {c['row']} column {c['col']} = {v!r}"
        if "approx" in c:
            val, tol = c["approx"]
            assert abs(v - val) < tol, where
        if "abs_lt" in c:
            assert abs(v) < c["abs_lt"], where
        if "eq" in c:
            assert v == c["eq"], where
