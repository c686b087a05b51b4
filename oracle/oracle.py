"""ctypes access to the CPU checkers (test infrastructure only).

  liboracle.so            the C restatement (simplify_oracle.c)
  _ref/libmeshopt_ref.so  the reference library compiled from its own sources (optional)

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmeshopt_ref.so")
REF_TRACE_SO = os.path.join(HERE, "_ref", "libmeshopt_ref_trace.so")

MAX_PASSES = 64


class PassStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint) for n in (
        "candidates", "evaluated", "committed", "skipped", "rejected",
        "rounds", "window", "stop", "tri_collapses", "live_triangles")]


class Stats(ctypes.Structure):
    _fields_ = [
        ("passes", ctypes.c_int),
        ("unique_positions", ctypes.c_uint),
        ("kinds", ctypes.c_uint * 5),
        ("result_error_sq", ctypes.c_float),
        ("vertex_error_sq", ctypes.c_float),
        ("solve_rounds", ctypes.c_uint),
        ("pass_", PassStats * MAX_PASSES),
    ]

    def pass_list(self):
        out = []
        for p in range(min(self.passes, MAX_PASSES)):
            s = self.pass_[p]
            out.append({f: getattr(s, f) for f, _ in PassStats._fields_})
        return out


class Config(ctypes.Structure):
    _fields_ = [("perform_rounds", ctypes.c_int), ("round_window", ctypes.c_size_t), ("solve_rounds", ctypes.c_int)]


_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_F = ctypes.c_float
_U = ctypes.c_uint

_SIG_SIMPLIFY = [_P, _P, _SZ, _P, _SZ, _SZ, _SZ, _F, _U, _P]
_SIG_ATTR = [_P, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _F, _U, _P]

_libs: dict[str, ctypes.CDLL] = {}


def build(ref_dir: str = "/root/reference") -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, f"REF={ref_dir}"], check=True)


def _load(path: str) -> ctypes.CDLL:
    lib = _libs.get(path)
    if lib is None:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        lib = ctypes.CDLL(path)
        _libs[path] = lib
    return lib


def oracle_lib() -> ctypes.CDLL:
    lib = _load(ORACLE_SO)
    if not getattr(lib, "_typed", False):
        lib.moracle_simplify_edge.restype = _SZ
        lib.moracle_simplify_edge.argtypes = [_P, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _F, _U, _P, _P, _P]
        lib.moracle_simplify_scale.restype = _F
        lib.moracle_simplify_scale.argtypes = [_P, _SZ, _SZ]
        lib._typed = True
    return lib


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref_lib(trace: bool = False) -> ctypes.CDLL:
    lib = _load(REF_TRACE_SO if trace else REF_SO)
    if not getattr(lib, "_typed", False):
        lib.meshopt_simplify.restype = _SZ
        lib.meshopt_simplify.argtypes = _SIG_SIMPLIFY
        lib.meshopt_simplifyWithAttributes.restype = _SZ
        lib.meshopt_simplifyWithAttributes.argtypes = _SIG_ATTR
        lib.meshopt_simplifyScale.restype = _F
        lib.meshopt_simplifyScale.argtypes = [_P, _SZ, _SZ]
        lib._typed = True
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _rows(a, name):
    """A float32 2-D array whose rows are vertices; may be a strided view into a wider buffer."""
    a = np.asarray(a)
    if a.dtype != np.float32:
        raise TypeError(f"{name} must be float32")
    if a.ndim == 1:
        a = a.reshape(-1, 3)
    if a.shape[0] == 0:
        # numpy gives empty arrays zero strides; keep a packed row stride for the C side
        return np.zeros((1, a.shape[1]), dtype=np.float32)[:0]
    if a.strides[1] != 4 or a.strides[0] % 4 or a.strides[0] < 4 * a.shape[1]:
        raise ValueError(f"{name} rows must be packed float32 with a row stride >= its width")
    return a


def simplify(indices, positions, target_index_count, target_error, options=0, *,
             attributes=None, weights=None, lock=None,
             rounds=False, window=0, stats=False, impl="oracle"):
    """Run a CPU simplifier.

    positions:  float32 [V, >=3] (row stride = vertex stride; may view an interleaved buffer)
    attributes: float32 [V, A] view (row stride = attribute stride) or None for meshopt_simplify
    impl: "oracle" (the C restatement) or "ref" / "ref_trace" (the compiled reference library).
    Returns (result_indices, result_error[, Stats]).
    """
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    positions = _rows(positions, "positions")
    vcount = positions.shape[0]
    dest = np.zeros(max(indices.size, 1), dtype=np.uint32)
    err = ctypes.c_float(0)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    lk = None if lock is None else np.ascontiguousarray(lock, dtype=np.uint8)
    acount = 0
    aptr = None
    astride = 0
    if attributes is not None and np.asarray(attributes).size:
        attributes = _rows(attributes, "attributes")
        acount = attributes.shape[1]
        aptr = attributes.ctypes.data
        astride = attributes.strides[0]

    st = Stats()
    if impl == "oracle":
        lib = oracle_lib()
        cfg = Config(1 if rounds else 0, window, 0)
        n = lib.moracle_simplify_edge(
            dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data, vcount, positions.strides[0],
            aptr, astride, _ptr(w), acount, _ptr(lk),
            target_index_count, target_error, options, ctypes.byref(err), ctypes.byref(cfg), ctypes.byref(st))
    elif impl in ("ref", "ref_trace"):
        lib = ref_lib(trace=impl == "ref_trace")
        if attributes is None and lock is None and weights is None:
            n = lib.meshopt_simplify(dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data,
                                     vcount, positions.strides[0], target_index_count, target_error, options,
                                     ctypes.byref(err))
        else:
            n = lib.meshopt_simplifyWithAttributes(
                dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data, vcount,
                positions.strides[0], aptr, astride, _ptr(w), acount, _ptr(lk),
                target_index_count, target_error, options, ctypes.byref(err))
    else:
        raise ValueError(impl)
    out = dest[:n].copy()
    if stats:
        return out, err.value, st
    return out, err.value


def _writable_rows(a, name):
    a = _rows(a, name)
    if a.size and not a.flags.writeable:
        raise ValueError(f"{name} must be writable (it is updated in place)")
    return a


SOLVE = 1 << 29  # the internal option bit meshopt_simplifyWithUpdate adds (src/simplifier.cpp:2340)


def simplify_with_update(indices, positions, target_index_count, target_error, options=0, *,
                         attributes=None, weights=None, lock=None, impl="oracle", solve_rounds=False, stats=False):
    """meshopt_simplifyWithUpdate on the CPU: the C restatement (impl="oracle") or the compiled
    reference (impl="ref").  positions [V, >=3] and attributes [V, A] are float32 row views (they
    may share one interleaved buffer) and are UPDATED IN PLACE, as the reference does.
    solve_rounds runs the oracle's position solve in dependency rounds (the device's schedule).
    Returns (result_indices, result_error[, Stats])."""
    ib = np.array(indices, dtype=np.uint32, copy=True).reshape(-1)
    if ib.size == 0:
        ib = np.zeros(1, dtype=np.uint32)[:0]
    positions = _writable_rows(positions, "positions")
    vcount = positions.shape[0]
    err = ctypes.c_float(0)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    lk = None if lock is None else np.ascontiguousarray(lock, dtype=np.uint8)
    acount, aptr, astride = 0, None, 0
    if attributes is not None and np.asarray(attributes).size:
        attributes = _writable_rows(attributes, "attributes")
        acount, aptr, astride = attributes.shape[1], attributes.ctypes.data, attributes.strides[0]
    args = (ib.ctypes.data, ib.size, positions.ctypes.data, vcount, positions.strides[0], aptr, astride, _ptr(w), acount,
            _ptr(lk), target_index_count, target_error, options, ctypes.byref(err))
    sig = [_P, _SZ, _P, _SZ, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _F, _U, _P]
    st = Stats()
    if impl == "oracle":
        cfg = Config(0, 0, 1 if solve_rounds else 0)
        n = oracle_lib().moracle_simplify_edge(
            ib.ctypes.data, ib.ctypes.data, ib.size, positions.ctypes.data, vcount, positions.strides[0], aptr, astride,
            _ptr(w), acount, _ptr(lk), target_index_count, target_error, options | SOLVE, ctypes.byref(err),
            ctypes.byref(cfg), ctypes.byref(st))
    elif impl == "ref":
        fn = ref_lib().meshopt_simplifyWithUpdate
        fn.restype = _SZ
        fn.argtypes = sig
        n = fn(*args)
    else:
        raise ValueError(impl)
    if stats:
        return ib[:n].copy(), err.value, st
    return ib[:n].copy(), err.value


class SloppyTrace(ctypes.Structure):
    _fields_ = [("passes", ctypes.c_int), ("final_grid", ctypes.c_int), ("cells", ctypes.c_uint),
                ("grid", ctypes.c_int * MAX_PASSES), ("triangles", ctypes.c_uint * MAX_PASSES)]


def simplify_sloppy(indices, positions, target_index_count, target_error, *, lock=None, impl="oracle", trace=False):
    """meshopt_simplifySloppy on the CPU: the C restatement (impl="oracle") or the compiled reference
    (impl="ref").  Returns (result_indices, result_error[, SloppyTrace])."""
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    positions = _rows(positions, "positions")
    vcount = positions.shape[0]
    dest = np.zeros(max(indices.size, 1), dtype=np.uint32)
    err = ctypes.c_float(0)
    lk = None if lock is None else np.ascontiguousarray(lock, dtype=np.uint8)
    tr = SloppyTrace()
    if impl == "oracle":
        lib = oracle_lib()
        fn = lib.moracle_simplify_sloppy
        fn.restype = _SZ
        fn.argtypes = [_P, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _F, _P, _P]
        n = fn(dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data, vcount, positions.strides[0],
               _ptr(lk), target_index_count, target_error, ctypes.byref(err), ctypes.byref(tr))
    elif impl == "ref":
        fn = ref_lib().meshopt_simplifySloppy
        fn.restype = _SZ
        fn.argtypes = [_P, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _F, _P]
        n = fn(dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data, vcount, positions.strides[0],
               _ptr(lk), target_index_count, target_error, ctypes.byref(err))
    else:
        raise ValueError(impl)
    out = dest[:n].copy()
    if trace:
        return out, err.value, tr
    return out, err.value


def simplify_scale(vertices, impl="oracle"):
    vertices = np.ascontiguousarray(vertices, dtype=np.float32)
    if vertices.ndim == 1:
        vertices = vertices.reshape(-1, 3)
    lib = oracle_lib() if impl == "oracle" else ref_lib()
    fn = lib.moracle_simplify_scale if impl == "oracle" else lib.meshopt_simplifyScale
    return fn(vertices.ctypes.data, vertices.shape[0], vertices.shape[1] * 4)


BATCH_SO = os.path.join(HERE, "libbatch_driver.so")


def simplify_batch(indices, index_offsets, positions, vertex_offsets, targets, target_error, options=0, threads=1,
                   impl="ref"):
    """Runs the reference library (impl="ref") or the C restatement (impl="oracle") over a packed
    batch of meshes on `threads` POSIX threads (batch_driver.c).  positions is [V, k] float32.
    Returns (destination, result_counts, result_errors)."""
    drv = _load(BATCH_SO)
    if impl == "ref":
        fn = ctypes.cast(ref_lib().meshopt_simplify, ctypes.c_void_p).value
    elif impl == "oracle":
        fn = ctypes.cast(oracle_lib().moracle_simplify, ctypes.c_void_p).value
    else:
        raise ValueError(impl)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    io = np.ascontiguousarray(index_offsets, dtype=np.uint32)
    vo = np.ascontiguousarray(vertex_offsets, dtype=np.uint32)
    tg = np.ascontiguousarray(targets, dtype=np.uint32)
    positions = np.ascontiguousarray(positions, dtype=np.float32)
    count = tg.size
    dest = np.zeros(max(indices.size, 1), dtype=np.uint32)
    counts = np.zeros(max(count, 1), dtype=np.uint32)
    errors = np.zeros(max(count, 1), dtype=np.float32)
    drv.batch_simplify.restype = ctypes.c_int
    drv.batch_simplify.argtypes = [_P, ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _SZ, _F, _U]
    rc = drv.batch_simplify(fn, int(threads), dest.ctypes.data, counts.ctypes.data, errors.ctypes.data,
                            indices.ctypes.data, io.ctypes.data, positions.ctypes.data, vo.ctypes.data,
                            positions.strides[0], tg.ctypes.data, count, float(target_error), int(options))
    if rc != 0:
        raise RuntimeError("batch_simplify could not start its threads")
    return dest, counts[:count], errors[:count]
