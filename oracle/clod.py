"""CPU restatement of clusterlod's per-group simplify (demo/clusterlod.h:426-518), test
infrastructure only: the group's Sparse call, the Permissive re-run, the sloppy fallback over the
deindexed triangles and the edge-length error limit, composed from the oracle's (or the compiled
reference's) meshopt_simplifyWithAttributes, meshopt_simplifySloppy and meshopt_simplifyScale.

Only tests/ and bench.py's CPU-baseline leg import this module.

Pinned by: the composed calls' own parity (the oracle against the reference's golden vectors and
the compiled reference); edge_limit_sq against closed forms (equilateral, wire and degenerate
triangles, a NaN vertex) in tests/test_oracle.py.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

FLT_MAX = float(np.finfo(np.float32).max)
SPARSE = 1 << 1
PERMISSIVE = 1 << 5

DEFAULT_CONFIG = {"fallback_permissive": 0, "fallback_sloppy": 1, "error_factor_sloppy": 2.0, "error_edge_limit": 0.0}


def _std_max(a, b):
    # std::max(a, b): b only when a < b (a NaN in b never wins; one in a stays)
    return np.where(a < b, b, a)


def _std_min(a, b):
    # std::min(a, b): b only when b < a
    return np.where(b < a, b, a)


def edge_limit_sq(indices, positions):
    """max over the triangles of std::max(shortest squared edge, longest / 4), the running maximum
    starting at 0 with std::max (demo/clusterlod.h:491-512), in float32."""
    p = np.ascontiguousarray(positions, dtype=np.float32)[:, :3]
    t = np.asarray(indices, dtype=np.int64).reshape(-1, 3)
    if t.shape[0] == 0:
        return np.float32(0)
    a, b, c = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]

    def sq(u, v):
        d = u - v
        return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]

    eab, eac, ebc = sq(a, b), sq(a, c), sq(b, c)
    emax = _std_max(_std_max(eab, eac), ebc)
    emin = _std_min(_std_min(eab, eac), ebc)
    cand = _std_max(emin, emax / np.float32(4))
    m = np.float32(0)
    for x in cand:  # the running std::max in triangle order
        if m < x:
            m = x
    return np.float32(m)


def clod_simplify(indices, positions, locks, target_count, options, config=None, impl="oracle", attributes=None, weights=None):
    """One group as clodBuild's simplify() returns it: (result indices in the shared vertex ids,
    float32 error).  options are the level's (Sparse is added, as the demo always passes it);
    config holds meshopt_hip_clod_config's fields (None: no fallbacks, no edge limit); attributes
    (float32 [V, A] of the shared vertices) and weights as the demo's clodMesh carries them (None:
    none)."""
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    positions = np.ascontiguousarray(positions, dtype=np.float32)
    cfg = config or {"fallback_permissive": 0, "fallback_sloppy": 0, "error_factor_sloppy": 1.0, "error_edge_limit": 0.0}
    if target_count > indices.size:
        return indices.copy(), np.float32(0)
    if attributes is None:
        attributes, weights = np.zeros((positions.shape[0], 0), np.float32), np.zeros(0, np.float32)
    else:
        attributes = np.ascontiguousarray(attributes, dtype=np.float32)
        weights = np.ascontiguousarray(weights, dtype=np.float32)
    lod, err = O.simplify(indices, positions, target_count, FLT_MAX, options | SPARSE, attributes=attributes, weights=weights, lock=locks,
                          impl=impl)
    err = np.float32(err)
    if lod.size > target_count and cfg["fallback_permissive"] and not (options & PERMISSIVE):
        lod, err = O.simplify(indices, positions, target_count, FLT_MAX, options | SPARSE | PERMISSIVE, attributes=attributes, weights=weights,
                              lock=locks, impl=impl)
        err = np.float32(err)
    if lod.size > target_count and cfg["fallback_sloppy"]:
        # simplifyFallback: one vertex per index, the sloppy call over the identity buffer
        subset = np.ascontiguousarray(positions[indices, :3])
        sub_lod, e = O.simplify_sloppy(np.arange(indices.size, dtype=np.uint32), subset, target_count, FLT_MAX,
                                       lock=np.ascontiguousarray(np.asarray(locks, np.uint8)[indices]), impl=impl)
        err = np.float32(e) * np.float32(O.simplify_scale(subset, impl=impl))
        err = np.float32(err * np.float32(cfg["error_factor_sloppy"]))
        lod = indices[sub_lod]
    limit = np.float32(cfg["error_edge_limit"])
    if limit > 0:
        bound = np.float32(np.sqrt(edge_limit_sq(indices, positions)) * limit)
        err = bound if bound < err else err
    return lod, np.float32(err)
