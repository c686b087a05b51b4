"""Runs the batched device code through a sanitizer build of the CPU harness (argv[1]: the built
library) on patches, random soups, seamed grids and edge-value meshes, each mesh checked against the
oracle; exits non-zero on a mismatch (the sanitizers abort the process on their own findings)."""
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np

import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W
from meshes import edge_value_case, random_grid, random_soup
import test_simt_batch as T


def main():
    lib = ctypes.CDLL(sys.argv[1])
    P, SZ = ctypes.c_void_p, ctypes.c_size_t
    lib.simt_simplify_batch.restype = ctypes.c_int
    lib.simt_simplify_batch.argtypes = [P, P, P, P, P, P, P, P, P, SZ, P, SZ, ctypes.c_float, ctypes.c_uint, ctypes.c_uint,
                                        ctypes.c_uint, ctypes.c_int, ctypes.c_uint, ctypes.c_ulonglong]
    lib.simt_retries.restype = ctypes.c_uint
    lib.simt_retries.argtypes = []
    lib.simt_set_vertex_lock.restype = None
    lib.simt_set_vertex_lock.argtypes = [P]
    ms = W.patches(count=16, seed=3)
    meshes = [(m.indices, np.ascontiguousarray(m.vertices[:, :3])) for m in ms]
    targets = [W.target_count(m.index_count, 1, 2) for m in ms]
    T.check(lib, meshes, targets, 1e-2, M.SimplifyLockBorder, seed=2)
    rng = np.random.default_rng(61)
    meshes, targets = [], []
    for k in range(8):
        if k % 2:
            ib, v = random_soup(rng, nv=int(rng.integers(3, 150)), ntri=int(rng.integers(1, 300)))
        else:
            ib, v = random_grid(rng, n=int(rng.integers(3, 16)), seams=int(rng.integers(0, 4)))
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    T.check(lib, meshes, targets, 1e-1, M.SimplifyPermissive, seed=5)
    meshes, targets = [], []
    for seed in range(8):
        ib, v, _ = edge_value_case(seed)
        meshes.append((ib, np.ascontiguousarray(v[:, :3])))
        targets.append((ib.size // 3 * int(np.random.default_rng(seed).integers(0, 4)) // 4) * 3)
    T.check(lib, meshes, targets, 1e-1, 0, seed=7)
    print("sanitized run ok", flush=True)


if __name__ == "__main__":
    main()
