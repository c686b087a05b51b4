"""meshoptimizer_amd -- MI355X (gfx950) implementation of meshoptimizer's quadric
edge-collapse simplifier behind the reference's own C interface.

The product is the shared library libmeshopt_hip.so (C ABI declared in include/
meshoptimizer.h and include/meshopt_hip.h).  This module is a thin Python view of that ABI
with the reference's names and argument meaning:

  simplify(indices, positions, target_index_count, target_error, options)
      -> meshopt_simplify (src/meshoptimizer.h:513)
  simplify_with_attributes(indices, positions, attributes, weights, lock, target_index_count,
                           target_error, options)
      -> meshopt_simplifyWithAttributes (src/meshoptimizer.h:537)
  simplify_with_update(indices, positions, target_index_count, target_error, options,
                       attributes, weights, lock)
      -> meshopt_simplifyWithUpdate (src/meshoptimizer.h:560); updates its arrays in place
  simplify_sloppy(indices, positions, target_index_count, target_error, lock)
      -> meshopt_simplifySloppy (src/meshoptimizer.h:576)
  simplify_scale(positions)
      -> meshopt_simplifyScale (src/meshoptimizer.h:611)
  simplify_device(...)  device-resident variant (buffers already in HBM)

There is no CPU fallback: if the HIP library is missing the import of the library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmeshopt_hip.so")

# option bits, src/meshoptimizer.h:466-482
SimplifyLockBorder = 1 << 0
SimplifySparse = 1 << 1
SimplifyErrorAbsolute = 1 << 2
SimplifyPrune = 1 << 3
SimplifyRegularize = 1 << 4
SimplifyPermissive = 1 << 5
SimplifyRegularizeLight = 1 << 6
SimplifyInternalDebug = 1 << 30

# vertex_lock flags, src/meshoptimizer.h:487-495
SimplifyVertex_Lock = 1 << 0
SimplifyVertex_Protect = 1 << 1
SimplifyVertex_Priority = 1 << 2

# every symbol the headers in include/ declare
EXPORTED_SYMBOLS = (
    "meshopt_simplify",
    "meshopt_simplifyWithAttributes",
    "meshopt_simplifyScale",
    "meshopt_simplifySloppy",
    "meshopt_simplifyWithUpdate",
    "meshopt_setAllocator",
    "meshopt_hip_simplify_with_update_device",
    "meshopt_hip_simplify_sloppy_device",
    "meshopt_hip_simplify_device",
    "meshopt_hip_simplify_batch",
    "meshopt_hip_simplify_batch_device",
    "meshopt_hip_simplify_batch_device_noalloc",
    "meshopt_hip_simplify_batch_with_attributes",
    "meshopt_hip_simplify_batch_with_attributes_device",
    "meshopt_hip_clod_level_device",
    "meshopt_hip_batch_kernel_timing",
    "meshopt_hip_last_batch_stats",
    "meshopt_hip_batch_set_shape",
    "meshopt_hip_batch_get_shape",
    "meshopt_hip_server_create",
    "meshopt_hip_server_submit",
    "meshopt_hip_server_wait",
    "meshopt_hip_server_clock_ms",
    "meshopt_hip_server_destroy",
    "meshopt_hip_last_call_stats",
    "meshopt_hip_set_round_policy",
    "meshopt_hip_reset_round_counts",
    "meshopt_hip_last_error",
    "meshopt_hip_kernel_timing",
    "meshopt_hip_release",
)


class PassStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint) for n in (
        "live_triangles", "candidates", "evaluated", "committed", "skipped", "rejected", "stop", "rounds", "tail_rounds")]


class CallStats(ctypes.Structure):
    _fields_ = [
        ("passes", ctypes.c_uint),
        ("vertex_count", ctypes.c_uint),
        ("unique_positions", ctypes.c_uint),
        ("kinds", ctypes.c_uint * 5),
        ("result_error_sq", ctypes.c_float),
        ("vertex_error_sq", ctypes.c_float),
        ("solve_vertices", ctypes.c_uint),
        ("solve_rounds", ctypes.c_uint),
        ("read_backs", ctypes.c_uint),
        ("device_extensions", ctypes.c_uint),
    ]


class ClodConfig(ctypes.Structure):
    """meshopt_hip_clod_config: clodConfig's fallback fields (demo/clusterlod.h:42-53)."""
    _fields_ = [
        ("fallback_permissive", ctypes.c_int),
        ("fallback_sloppy", ctypes.c_int),
        ("error_factor_sloppy", ctypes.c_float),
        ("error_edge_limit", ctypes.c_float),
    ]


_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_F = ctypes.c_float
_U = ctypes.c_uint
_lib = None


def library() -> ctypes.CDLL:
    """Load libmeshopt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m meshoptimizer_amd.build` "
                          "(the simplifier has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.meshopt_simplify.restype = _SZ
    lib.meshopt_simplify.argtypes = [_P, _P, _SZ, _P, _SZ, _SZ, _SZ, _F, _U, _P]
    lib.meshopt_simplifyWithAttributes.restype = _SZ
    lib.meshopt_simplifyWithAttributes.argtypes = [_P, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _F, _U, _P]
    lib.meshopt_simplifyScale.restype = _F
    lib.meshopt_simplifyScale.argtypes = [_P, _SZ, _SZ]
    for name in ("meshopt_simplifySloppy", "meshopt_hip_simplify_sloppy_device"):
        fn = getattr(lib, name)
        fn.restype = _SZ
        fn.argtypes = [_P, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _F, _P]
    for name in ("meshopt_simplifyWithUpdate", "meshopt_hip_simplify_with_update_device"):
        fn = getattr(lib, name)
        fn.restype = _SZ
        fn.argtypes = [_P, _SZ, _P, _SZ, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _F, _U, _P]
    lib.meshopt_hip_simplify_device.restype = _SZ
    lib.meshopt_hip_simplify_device.argtypes = [_P, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _F, _U, _P]
    lib.meshopt_hip_last_call_stats.restype = ctypes.c_int
    lib.meshopt_hip_last_call_stats.argtypes = [_P, _P, ctypes.c_int]
    lib.meshopt_hip_last_error.restype = ctypes.c_char_p
    lib.meshopt_hip_last_error.argtypes = []
    lib.meshopt_hip_simplify_batch.restype = ctypes.c_int
    lib.meshopt_hip_simplify_batch.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _SZ, _F, _U]
    lib.meshopt_hip_simplify_batch_with_attributes.restype = ctypes.c_int
    lib.meshopt_hip_simplify_batch_with_attributes.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _SZ, _P, _SZ, _P, _SZ, _F, _U]
    lib.meshopt_hip_simplify_batch_with_attributes_device.restype = ctypes.c_int
    lib.meshopt_hip_simplify_batch_with_attributes_device.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _F, _U]
    lib.meshopt_hip_clod_level_device.restype = ctypes.c_int
    lib.meshopt_hip_clod_level_device.argtypes = [_P, _P, _P, _P, _P, _SZ, _P, _SZ, _SZ, _P, _SZ, _P, _SZ, _P, _P, _P, _P, _F, _U,
                                                  ctypes.POINTER(ClodConfig)]
    lib.meshopt_hip_simplify_batch_device.restype = ctypes.c_int
    lib.meshopt_hip_simplify_batch_device.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _SZ, _P, _SZ, _F, _U]
    lib.meshopt_hip_simplify_batch_device_noalloc.restype = ctypes.c_int
    lib.meshopt_hip_simplify_batch_device_noalloc.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _SZ, _SZ, _P, _SZ, _F, _U]
    lib.meshopt_hip_last_batch_stats.restype = ctypes.c_int
    lib.meshopt_hip_batch_set_shape.restype = ctypes.c_int
    lib.meshopt_hip_batch_set_shape.argtypes = [_P, _P, _P]
    lib.meshopt_hip_batch_get_shape.restype = ctypes.c_int
    lib.meshopt_hip_batch_get_shape.argtypes = [_P, _P, _P, _P, _P]
    lib.meshopt_hip_last_batch_stats.argtypes = [_P, _SZ]
    lib.meshopt_hip_set_round_policy.restype = ctypes.c_int
    lib.meshopt_hip_set_round_policy.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.meshopt_hip_reset_round_counts.restype = ctypes.c_int
    lib.meshopt_hip_reset_round_counts.argtypes = []
    lib.meshopt_hip_batch_kernel_timing.restype = None
    lib.meshopt_hip_batch_kernel_timing.argtypes = [_P, _P, _P]
    lib.meshopt_hip_kernel_timing.restype = None
    lib.meshopt_hip_kernel_timing.argtypes = [ctypes.c_int, _P, _P, _P]
    lib.meshopt_hip_release.restype = None
    lib.meshopt_hip_release.argtypes = []
    lib.meshopt_hip_server_create.restype = _P
    lib.meshopt_hip_server_create.argtypes = [_SZ, ctypes.c_int]
    lib.meshopt_hip_server_submit.restype = _SZ
    lib.meshopt_hip_server_submit.argtypes = [_P, _P]
    lib.meshopt_hip_server_wait.restype = ctypes.c_int
    lib.meshopt_hip_server_wait.argtypes = [_P, _SZ, _P]
    lib.meshopt_hip_server_clock_ms.restype = ctypes.c_double
    lib.meshopt_hip_server_clock_ms.argtypes = [_P]
    lib.meshopt_hip_server_destroy.restype = None
    lib.meshopt_hip_server_destroy.argtypes = [_P]
    _lib = lib
    return lib


def _rows(a, name):
    a = np.asarray(a)
    if a.dtype != np.float32:
        raise TypeError(f"{name} must be float32")
    if a.ndim == 1:
        a = a.reshape(-1, 3)
    if a.shape[0] == 0:
        # numpy gives empty arrays zero strides; keep a packed row stride for the C side
        return np.zeros((1, a.shape[1]), dtype=np.float32)[:0]
    if a.shape[1] and (a.strides[1] != 4 or a.strides[0] % 4 or a.strides[0] < 4 * a.shape[1]):
        raise ValueError(f"{name} rows must be packed float32 with a row stride >= their width")
    return a


def simplify(indices, positions, target_index_count, target_error=1e-2, options=0):
    """meshopt_simplify.  positions: float32 [V, >=3] (row stride = vertex stride).
    Returns (indices uint32 [n], result_error)."""
    lib = library()
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    positions = _rows(positions, "positions")
    dest = np.zeros(max(indices.size, 1), dtype=np.uint32)
    err = ctypes.c_float(0)
    n = lib.meshopt_simplify(dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data,
                             positions.shape[0], positions.strides[0], int(target_index_count), float(target_error),
                             int(options), ctypes.byref(err))
    return dest[:n].copy(), err.value


def simplify_with_attributes(indices, positions, attributes, weights, lock, target_index_count,
                             target_error=1e-2, options=0):
    """meshopt_simplifyWithAttributes.  attributes: float32 [V, A] view or None; weights [A];
    lock: uint8 [V] or None.  Returns (indices, result_error)."""
    lib = library()
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    positions = _rows(positions, "positions")
    aptr, astride, acount = None, 0, 0
    if attributes is not None and np.asarray(attributes).size:
        attributes = _rows(attributes, "attributes")
        aptr, astride, acount = attributes.ctypes.data, attributes.strides[0], attributes.shape[1]
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    lk = None if lock is None else np.ascontiguousarray(lock, dtype=np.uint8)
    dest = np.zeros(max(indices.size, 1), dtype=np.uint32)
    err = ctypes.c_float(0)
    n = lib.meshopt_simplifyWithAttributes(
        dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data, positions.shape[0],
        positions.strides[0], aptr, astride, None if w is None else w.ctypes.data, acount,
        None if lk is None else lk.ctypes.data, int(target_index_count), float(target_error), int(options),
        ctypes.byref(err))
    return dest[:n].copy(), err.value


def simplify_with_update(indices, positions, target_index_count, target_error=1e-2, options=0, *,
                         attributes=None, weights=None, lock=None):
    """meshopt_simplifyWithUpdate.  positions: float32 [V, >=3] row view and attributes: float32
    [V, A] row view or None (they may view one interleaved buffer); both are UPDATED IN PLACE, as
    the reference updates its arrays.  weights [A]; lock: uint8 [V] or None.
    Returns (indices uint32 [n], result_error)."""
    lib = library()
    ib = np.array(indices, dtype=np.uint32, copy=True).reshape(-1)
    positions = _rows(positions, "positions")
    if positions.size and not positions.flags.writeable:
        raise ValueError("positions must be writable (they are updated in place)")
    aptr, astride, acount = None, 0, 0
    if attributes is not None and np.asarray(attributes).size:
        attributes = _rows(attributes, "attributes")
        if not attributes.flags.writeable:
            raise ValueError("attributes must be writable (they are updated in place)")
        aptr, astride, acount = attributes.ctypes.data, attributes.strides[0], attributes.shape[1]
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    lk = None if lock is None else np.ascontiguousarray(lock, dtype=np.uint8)
    err = ctypes.c_float(0)
    n = lib.meshopt_simplifyWithUpdate(
        ib.ctypes.data if ib.size else None, ib.size, positions.ctypes.data, positions.shape[0], positions.strides[0],
        aptr, astride, None if w is None else w.ctypes.data, acount, None if lk is None else lk.ctypes.data,
        int(target_index_count), float(target_error), int(options), ctypes.byref(err))
    return ib[:n].copy(), err.value


def simplify_with_update_device(indices_ptr, index_count, positions_ptr, vertex_count, position_stride,
                                target_index_count, target_error, options=0, attributes_ptr=None, attribute_stride=0,
                                weights=None, lock_ptr=None):
    """meshopt_hip_simplify_with_update_device on device pointers (integers): indices, positions and
    attributes are updated in place on the device.  Returns (index count, result_error); raises on error."""
    lib = library()
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    acount = 0 if w is None else w.size
    err = ctypes.c_float(0)
    n = lib.meshopt_hip_simplify_with_update_device(
        indices_ptr, index_count, positions_ptr, vertex_count, position_stride, attributes_ptr, attribute_stride,
        None if w is None else w.ctypes.data, acount, lock_ptr, int(target_index_count), float(target_error), int(options),
        ctypes.byref(err))
    if n == ctypes.c_size_t(-1).value:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())
    return n, err.value


def simplify_sloppy(indices, positions, target_index_count, target_error=1e-2, lock=None):
    """meshopt_simplifySloppy.  positions: float32 [V, >=3] (row stride = vertex stride); lock: uint8 [V]
    or None.  Returns (indices uint32 [n], result_error)."""
    lib = library()
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    positions = _rows(positions, "positions")
    lk = None if lock is None else np.ascontiguousarray(lock, dtype=np.uint8)
    dest = np.zeros(max(indices.size, 1), dtype=np.uint32)
    err = ctypes.c_float(0)
    n = lib.meshopt_simplifySloppy(dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data,
                                   positions.shape[0], positions.strides[0], None if lk is None else lk.ctypes.data,
                                   int(target_index_count), float(target_error), ctypes.byref(err))
    return dest[:n].copy(), err.value


def simplify_sloppy_device(destination_ptr, indices_ptr, index_count, positions_ptr, vertex_count, stride, lock_ptr,
                           target_index_count, target_error):
    """meshopt_hip_simplify_sloppy_device on device pointers (integers).  Returns (index count,
    result_error); raises on error."""
    lib = library()
    err = ctypes.c_float(0)
    n = lib.meshopt_hip_simplify_sloppy_device(destination_ptr, indices_ptr, index_count, positions_ptr, vertex_count, stride,
                                               lock_ptr, int(target_index_count), float(target_error), ctypes.byref(err))
    if n == ctypes.c_size_t(-1).value:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())
    return n, err.value


def simplify_scale(positions):
    """meshopt_simplifyScale."""
    lib = library()
    positions = _rows(positions, "positions")
    return lib.meshopt_simplifyScale(positions.ctypes.data, positions.shape[0], positions.strides[0])


def simplify_device(dest_ptr, indices_ptr, index_count, positions_ptr, vertex_count, position_stride,
                    target_index_count, target_error, options=0, attributes_ptr=None, attribute_stride=0,
                    weights=None, lock_ptr=None):
    """meshopt_hip_simplify_device on device pointers (e.g. torch tensor data_ptr()).
    Returns (index_count, result_error)."""
    lib = library()
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    acount = 0 if w is None else w.size
    err = ctypes.c_float(0)
    n = lib.meshopt_hip_simplify_device(
        dest_ptr, indices_ptr, index_count, positions_ptr, vertex_count, position_stride,
        attributes_ptr, attribute_stride, None if w is None else w.ctypes.data, acount, lock_ptr,
        int(target_index_count), float(target_error), int(options), ctypes.byref(err))
    if n == ctypes.c_size_t(-1).value:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())
    return n, err.value


def pack_meshes(meshes):
    """Packs a list of (indices, positions[V, >=3] float32) into batch arrays:
    (indices, index_offsets, vertices, vertex_offsets, stride_bytes)."""
    io = np.zeros(len(meshes) + 1, dtype=np.uint32)
    vo = np.zeros(len(meshes) + 1, dtype=np.uint32)
    for p, (ib, v) in enumerate(meshes):
        io[p + 1] = io[p] + len(ib)
        vo[p + 1] = vo[p] + len(v)
    width = max((np.asarray(v).shape[1] for _, v in meshes), default=3)
    verts = np.zeros((int(vo[-1]), width), dtype=np.float32)
    idx = np.zeros(int(io[-1]), dtype=np.uint32)
    for p, (ib, v) in enumerate(meshes):
        idx[io[p]:io[p + 1]] = ib
        verts[vo[p]:vo[p + 1], :np.asarray(v).shape[1]] = v
    return idx, io, verts, vo, width * 4


def simplify_batch(meshes, target_index_counts, target_error=1e-2, options=0, locks=None, attributes=None, weights=None):
    """Simplifies every (indices, positions) mesh in one batched launch (meshopt_simplify per mesh,
    or meshopt_simplifyWithAttributes with the mesh's vertex_lock when `locks` gives one uint8
    array per mesh, and its attributes when `attributes` gives one float32 [V, A] array per mesh
    with the shared `weights`).  Returns [(indices, result_error)]."""
    lib = library()
    idx, io, verts, vo, stride = pack_meshes(meshes)
    lk = None
    if locks is not None:
        lk = np.zeros(max(int(vo[-1]), 1), dtype=np.uint8)
        for p, l in enumerate(locks):
            lk[vo[p]:vo[p + 1]] = np.asarray(l, dtype=np.uint8)
    tg = np.ascontiguousarray(target_index_counts, dtype=np.uint32)
    dest = np.zeros(max(idx.size, 1), dtype=np.uint32)
    counts = np.zeros(max(len(meshes), 1), dtype=np.uint32)
    errors = np.zeros(max(len(meshes), 1), dtype=np.float32)
    if attributes is not None:
        w = np.ascontiguousarray(weights, dtype=np.float32)
        at = np.zeros((max(int(vo[-1]), 1), w.size), dtype=np.float32)
        for p, a in enumerate(attributes):
            at[vo[p]:vo[p + 1]] = np.asarray(a, dtype=np.float32).reshape(-1, w.size)
        rc = lib.meshopt_hip_simplify_batch_with_attributes(dest.ctypes.data, counts.ctypes.data, errors.ctypes.data, idx.ctypes.data,
                                                            io.ctypes.data, verts.ctypes.data, vo.ctypes.data, at.ctypes.data, 4 * w.size,
                                                            w.ctypes.data, w.size, None if lk is None else lk.ctypes.data, stride,
                                                            tg.ctypes.data, len(meshes), float(target_error), int(options))
    else:
        rc = lib.meshopt_hip_simplify_batch(dest.ctypes.data, counts.ctypes.data, errors.ctypes.data, idx.ctypes.data, io.ctypes.data,
                                            verts.ctypes.data, vo.ctypes.data, None if lk is None else lk.ctypes.data, stride,
                                            tg.ctypes.data, len(meshes),
                                            float(target_error), int(options))
    if rc != 0:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())
    return [(dest[io[p]:io[p] + counts[p]].copy(), float(errors[p])) for p in range(len(meshes))]


def kernel_timing(enable: bool):
    """Turns HIP-event timing of the round kernel (k_round_x) on/off; returns (ms, launches, ranks)
    gathered since the previous call."""
    lib = library()
    ms = ctypes.c_double(0)
    n = ctypes.c_size_t(0)
    r = ctypes.c_size_t(0)
    lib.meshopt_hip_kernel_timing(1 if enable else 0, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(r))
    return ms.value, n.value, r.value


def batch_kernel_timing():
    """(ms, launches, meshes) of the batched kernel gathered since the previous call, while
    kernel_timing(True) is on."""
    lib = library()
    ms = ctypes.c_double(0)
    n = ctypes.c_size_t(0)
    k = ctypes.c_size_t(0)
    lib.meshopt_hip_batch_kernel_timing(ctypes.byref(ms), ctypes.byref(n), ctypes.byref(k))
    return ms.value, n.value, k.value


def batch_set_shape(per_cu=None, threads=None, cursor=None):
    """Sets the launch shape of the batched calls' four size classes (workgroups per CU, threads per
    workgroup, cursor form); no arguments restore the default.  Raises with the reason on an invalid
    shape."""
    lib = library()
    if per_cu is None and threads is None and cursor is None:
        rc = lib.meshopt_hip_batch_set_shape(None, None, None)
    else:
        arrs = [np.ascontiguousarray(x, dtype=np.uint32) for x in (per_cu, threads, cursor)]
        if any(a.size != 4 for a in arrs):
            raise ValueError("a shape gives four values per field, one per size class")
        rc = lib.meshopt_hip_batch_set_shape(arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data)
    if rc != 0:
        raise ValueError(lib.meshopt_hip_last_error().decode())


def batch_get_shape():
    """The launch shape in use and the vertex and index caps it gives each size class on this
    device: dict of per_cu, threads, cursor, vertex_caps, index_caps (four entries each)."""
    lib = library()
    out = [np.zeros(4, dtype=np.uint32) for _ in range(5)]
    if lib.meshopt_hip_batch_get_shape(*(a.ctypes.data for a in out)) != 0:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())
    return dict(zip(("per_cu", "threads", "cursor", "vertex_caps", "index_caps"), (a.tolist() for a in out)))


def last_batch_stats(mesh_count):
    """Per-mesh statistics of the last batched call: a (n, 4) uint32 array of passes and, summed over
    the passes, live triangles at the pass start, candidates and committed collapses."""
    lib = library()
    out = np.zeros((max(mesh_count, 1), 4), dtype=np.uint32)
    n = lib.meshopt_hip_last_batch_stats(out.ctypes.data, mesh_count)
    if n < 0:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())
    return out[:n]


def simplify_batch_device(destination_ptr, result_counts_ptr, result_errors_ptr, indices_ptr, index_offsets_ptr,
                          positions_ptr, vertex_offsets_ptr, vertex_count, stride, target_counts_ptr, mesh_count,
                          target_error, options=0, lock_ptr=None):
    """meshopt_hip_simplify_batch_device on device pointers (integers); vertex_count is the number of
    vertices the positions hold, lock_ptr one vertex_lock byte per packed vertex or None.  Raises on
    error."""
    lib = library()
    rc = lib.meshopt_hip_simplify_batch_device(destination_ptr, result_counts_ptr, result_errors_ptr, indices_ptr,
                                               index_offsets_ptr, positions_ptr, vertex_offsets_ptr, lock_ptr, vertex_count,
                                               stride, target_counts_ptr, mesh_count, float(target_error), int(options))
    if rc != 0:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())


def simplify_batch_with_attributes_device(destination_ptr, result_counts_ptr, result_errors_ptr, indices_ptr, index_offsets_ptr,
                                          positions_ptr, vertex_offsets_ptr, attributes_ptr, attribute_stride, weights, vertex_count, stride,
                                          target_counts_ptr, mesh_count, target_error, options=0, lock_ptr=None):
    """meshopt_hip_simplify_batch_with_attributes_device on device pointers (integers); weights is a
    host float32 array (one weight per attribute, attribute_stride bytes per packed vertex).  Raises
    on error."""
    lib = library()
    w = np.ascontiguousarray(weights, dtype=np.float32)
    rc = lib.meshopt_hip_simplify_batch_with_attributes_device(destination_ptr, result_counts_ptr, result_errors_ptr, indices_ptr,
                                                               index_offsets_ptr, positions_ptr, vertex_offsets_ptr, attributes_ptr,
                                                               attribute_stride, w.ctypes.data, w.size, lock_ptr, vertex_count, stride,
                                                               target_counts_ptr, mesh_count, float(target_error), int(options))
    if rc != 0:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())


def simplify_batch_device_noalloc(destination_ptr, result_counts_ptr, result_errors_ptr, indices_ptr, index_offsets_ptr,
                                  positions_ptr, vertex_offsets_ptr, vertex_count, stride, target_counts_ptr, mesh_count,
                                  target_error, options=0, lock_ptr=None):
    """meshopt_hip_simplify_batch_device_noalloc (the batched device entry with no allocation, its
    scratch a static device array) on device pointers (integers).  Raises on error."""
    lib = library()
    rc = lib.meshopt_hip_simplify_batch_device_noalloc(destination_ptr, result_counts_ptr, result_errors_ptr, indices_ptr,
                                                       index_offsets_ptr, positions_ptr, vertex_offsets_ptr, lock_ptr, vertex_count,
                                                       stride, target_counts_ptr, mesh_count, float(target_error), int(options))
    if rc != 0:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())


def clod_level_device(destination_ptr, result_counts_ptr, result_errors_ptr, indices_ptr, index_offsets_ptr, group_count,
                      positions_ptr, vertex_count, stride, remap_ptr, base_locks_ptr, locks_ptr, target_counts_ptr,
                      target_error, options, config=None, attributes_ptr=None, attribute_stride=0, weights=None):
    """meshopt_hip_clod_level_device on device pointers (integers).  config: None (no fallbacks) or
    a dict with the keys of meshopt_hip_clod_config (fallback_permissive, fallback_sloppy,
    error_factor_sloppy, error_edge_limit).  attributes_ptr: the shared vertices' attributes
    (attribute_stride bytes each) with the host float32 array `weights`, or None.  Raises on
    error."""
    lib = library()
    w = np.ascontiguousarray(weights if weights is not None else np.zeros(0), dtype=np.float32)
    if attributes_ptr is None:
        w = np.zeros(0, dtype=np.float32)
    cfg = None
    if config is not None:
        cfg = ClodConfig(int(bool(config.get("fallback_permissive", 0))), int(bool(config.get("fallback_sloppy", 0))),
                         float(config.get("error_factor_sloppy", 1.0)), float(config.get("error_edge_limit", 0.0)))
    rc = lib.meshopt_hip_clod_level_device(destination_ptr, result_counts_ptr, result_errors_ptr, indices_ptr, index_offsets_ptr,
                                           group_count, positions_ptr, vertex_count, stride, attributes_ptr, attribute_stride,
                                           w.ctypes.data if w.size else None, w.size, remap_ptr, base_locks_ptr, locks_ptr,
                                           target_counts_ptr, float(target_error), int(options),
                                           ctypes.byref(cfg) if cfg is not None else None)
    if rc != 0:
        raise RuntimeError(lib.meshopt_hip_last_error().decode())


def set_round_policy(deep_heads=-2, deep_from_head=8):
    """meshopt_hip_set_round_policy: when the plain path's commit rounds examine the alternatives of
    moving corners (deep_heads -1: always; below -1: the default).  Raises on error."""
    lib = library()
    if lib.meshopt_hip_set_round_policy(int(deep_heads), int(deep_from_head)) != 0:
        raise ValueError(lib.meshopt_hip_last_error().decode())


def reset_round_counts():
    """meshopt_hip_reset_round_counts: forget the head-round counts learned per window size."""
    library().meshopt_hip_reset_round_counts()


def last_call_stats():
    """(call stats dict, [pass stats dict]) of the calling thread's last simplify call."""
    lib = library()
    call = CallStats()
    passes = (PassStats * 64)()
    n = lib.meshopt_hip_last_call_stats(ctypes.byref(call), passes, 64)
    cd = {f: getattr(call, f) for f, _ in CallStats._fields_}
    cd["kinds"] = list(call.kinds)
    return cd, [{f: getattr(passes[i], f) for f, _ in PassStats._fields_} for i in range(min(n, 64))]


class ServerRequest(ctypes.Structure):
    """meshopt_hip_request."""
    _fields_ = [("destination", _P), ("indices", _P), ("index_count", _SZ), ("vertex_positions", _P),
                ("vertex_count", _SZ), ("vertex_positions_stride", _SZ), ("vertex_attributes", _P),
                ("vertex_attributes_stride", _SZ), ("attribute_weights", _P), ("attribute_count", _SZ),
                ("vertex_lock", _P), ("target_index_count", _SZ), ("target_error", _F), ("options", _U)]


class ServerResponse(ctypes.Structure):
    """meshopt_hip_response."""
    _fields_ = [("index_count", _SZ), ("result_error", _F), ("slot", ctypes.c_int), ("submit_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("finish_ms", ctypes.c_double), ("device_bytes", _SZ),
                ("batched", ctypes.c_int)]


class Server:
    """A simplify server (meshopt_hip_server_*): requests run concurrently on `slots` device
    contexts and are admitted in arrival order against `budget_bytes` of device memory; small
    compatible requests queued together run as one batched call.  Each result equals simplify /
    simplify_with_attributes on the request alone."""

    def __init__(self, budget_bytes=64 << 30, slots=4):
        self._lib = library()
        self._h = self._lib.meshopt_hip_server_create(int(budget_bytes), int(slots))
        if not self._h:
            raise RuntimeError(self._lib.meshopt_hip_last_error().decode())
        self._live = {}

    def submit(self, indices, positions, target_index_count, target_error=1e-2, options=0, attributes=None,
               weights=None, lock=None):
        """Queues one request; returns its ticket.  The arrays are kept alive until wait()."""
        indices = np.ascontiguousarray(indices, dtype=np.uint32)
        positions = _rows(positions, "positions")
        aptr, astride, acount = None, 0, 0
        if attributes is not None and np.asarray(attributes).size:
            attributes = _rows(attributes, "attributes")
            aptr, astride, acount = attributes.ctypes.data, attributes.strides[0], attributes.shape[1]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        lk = None if lock is None else np.ascontiguousarray(lock, dtype=np.uint8)
        dest = np.zeros(max(indices.size, 1), dtype=np.uint32)
        req = ServerRequest(dest.ctypes.data, indices.ctypes.data, indices.size, positions.ctypes.data,
                            positions.shape[0], positions.strides[0], aptr, astride,
                            None if w is None else w.ctypes.data, acount if w is not None else 0,
                            None if lk is None else lk.ctypes.data, int(target_index_count), float(target_error),
                            int(options))
        ticket = self._lib.meshopt_hip_server_submit(self._h, ctypes.byref(req))
        if ticket == 0:
            raise RuntimeError(self._lib.meshopt_hip_last_error().decode())
        self._live[ticket] = (dest, indices, positions, attributes, w, lk)
        return ticket

    def wait(self, ticket):
        """Blocks until the request is done; returns (indices, result_error, info)."""
        resp = ServerResponse()
        rc = self._lib.meshopt_hip_server_wait(self._h, ticket, ctypes.byref(resp))
        dest = self._live.pop(ticket)[0]
        if rc != 0:
            raise RuntimeError(self._lib.meshopt_hip_last_error().decode())
        info = {"slot": resp.slot, "submit_ms": resp.submit_ms, "start_ms": resp.start_ms,
                "finish_ms": resp.finish_ms, "device_bytes": resp.device_bytes, "batched": resp.batched}
        return dest[:resp.index_count].copy(), resp.result_error, info

    def clock_ms(self):
        return self._lib.meshopt_hip_server_clock_ms(self._h)

    def close(self):
        if self._h:
            self._lib.meshopt_hip_server_destroy(self._h)
            self._h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
