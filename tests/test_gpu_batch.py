"""GPU parity of the batched entry point (meshopt_hip_simplify_batch): every mesh of a batch gets
exactly what the oracle's meshopt_simplify returns for it alone (index buffer, index count and
result_error bits)."""
import sys

import numpy as np
import pytest

from meshes import random_grid, random_soup, same_bits, scattered_components
from oracle import oracle as O
import meshoptimizer_amd as M
from meshoptimizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def announce(what):
    print(f"[case] {what}", flush=True)
    sys.stdout.flush()


def check_batch(meshes, targets, err, opts):
    got = M.simplify_batch(meshes, targets, err, opts)
    assert len(got) == len(meshes)
    for p, ((ib, v), t, (rg, eg)) in enumerate(zip(meshes, targets, got)):
        ro, eo = O.simplify(ib, v, int(t), err, opts)
        assert rg.size == ro.size, f"mesh {p}: index count {rg.size} != oracle {ro.size}"
        assert np.array_equal(rg, ro), f"mesh {p}: first difference at {int(np.argmax(rg != ro))}"
        assert same_bits(eg, eo), f"mesh {p}: result_error {eg!r} != oracle {eo!r}"
    return got


def patch_meshes(count, seed=4):
    ms = W.patches(count=count, seed=seed)
    return [(m.indices, m.vertices) for m in ms], [W.target_count(m.index_count, 1, 2) for m in ms]


def test_clusterlod_patches():
    """The batched workload's own shape: grid patches of 64-512 triangles, LockBorder, 50%."""
    announce("patches count=96 LockBorder target 50% err 1e-2")
    meshes, targets = patch_meshes(96)
    check_batch(meshes, targets, 1e-2, M.SimplifyLockBorder)


@pytest.mark.parametrize("opts", [0, M.SimplifyErrorAbsolute, M.SimplifyRegularize, M.SimplifyPermissive,
                                  M.SimplifyRegularizeLight, M.SimplifyLockBorder | M.SimplifyInternalDebug])
def test_patches_options(opts):
    announce(f"patches count=32 opts={opts}")
    meshes, targets = patch_meshes(32, seed=11)
    check_batch(meshes, targets, 5e-2, opts)


OPTION_SETS = [0, 1, 4, 16, 32, 64, 1 | 32, 32 | 16, 1 | 4, 1 << 30]


@pytest.mark.parametrize("seed", range(12))
def test_random_batches(seed):
    """Mixed random soups (repeated positions, degenerate triangles) and seamed grids in one batch."""
    rng = np.random.default_rng(7000 + seed)
    opts = OPTION_SETS[seed % len(OPTION_SETS)]
    err = float(rng.choice([1e-3, 1e-2, 1e-1, 1.0]))
    meshes, targets = [], []
    for k in range(int(rng.integers(4, 24))):
        if k % 2:
            ib, v = random_soup(rng, nv=int(rng.integers(3, 150)), ntri=int(rng.integers(1, 300)))
        else:
            ib, v = random_grid(rng, n=int(rng.integers(3, 16)))
        meshes.append((ib, v))
        targets.append((ib.size // 3 * int(rng.integers(0, 4)) // 4) * 3)
    announce(f"random batch seed={seed} meshes={len(meshes)} opts={opts} err={err}")
    check_batch(meshes, targets, err, opts)


def test_oversized_mesh_takes_plain_path():
    """A mesh too large for one workgroup's LDS is simplified by the plain path inside the same
    call; the small meshes around it still run in the batched kernel."""
    big = W.grid(160)
    meshes, targets = patch_meshes(12, seed=21)
    meshes.insert(5, (big.indices, big.vertices))
    targets.insert(5, W.target_count(big.index_count, 1, 4))
    announce(f"batch with one mesh of {big.index_count // 3} triangles")
    check_batch(meshes, targets, 1e-2, M.SimplifyLockBorder)


def test_sparse_batch():
    rng = np.random.default_rng(99)
    meshes, targets = [], []
    for _ in range(6):
        ib, v = random_soup(rng, nv=200, ntri=60)
        meshes.append((ib, v))
        targets.append(ib.size // 2 // 3 * 3)
    announce("sparse batch")
    check_batch(meshes, targets, 1e-1, M.SimplifySparse)


def test_empty_and_trivial_meshes():
    announce("empty batch, empty mesh, target above the count")
    assert M.simplify_batch([], [], 1e-2, 0) == []
    m = W.grid(6)
    empty = (np.zeros(0, np.uint32), np.zeros((0, 3), np.float32))
    meshes = [empty, (m.indices, m.vertices), (m.indices, m.vertices), empty]
    targets = [0, m.index_count, 0, 0]
    check_batch(meshes, targets, 1e-2, 0)


def test_bad_index_raises():
    announce("index out of range")
    m = W.grid(5)
    ib = m.indices.copy()
    ib[7] = m.vertex_count + 3
    with pytest.raises(RuntimeError, match="vertex_count"):
        M.simplify_batch([(m.indices, m.vertices), (ib, m.vertices)], [0, 0], 1e-2, 0)


def test_prune_batch():
    """Prune batches run mesh by mesh on the plain path inside the call, bit-exact per mesh, on
    meshes of many components at very different scales (Prune removes the small ones)."""
    rng = np.random.default_rng(31)
    meshes, targets = [], []
    for k in range(5):
        ib, v = scattered_components(rng, int(rng.integers(4, 16)))
        meshes.append((ib, v))
        targets.append(0 if k % 2 else (ib.size // 3 // 2) * 3)
    announce("prune batch")
    check_batch(meshes, targets, 5e-2, M.SimplifyPrune)
