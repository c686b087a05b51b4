"""Deterministic synthetic meshes for the simplifier workloads.

Every generator is seeded and built with numpy in float64 before a single rounding to
float32, so the same call gives the same buffers on any x86 host (the sine in the
icosphere displacement is the only libm call; a 1-ulp float64 difference almost never
survives the float32 rounding, and the CPU and GPU paths are always fed the very same
buffers in one process anyway).

Workloads (sizes match the simplifier benchmarks):
  grid(256)            256x256 grid, value-noise height 0.1      65,536 V / 130,050 T
  icosphere(8)         icosahedron subdivided 8 times, 8-sine
                       radial displacement 0.05                  655,362 V / 1,310,720 T
  seamed_grid(1024)    1024x1024 grid + normal/UV attributes with
                       16 UV seams (duplicated vertex lines)      ~1.06 M V / 2,093,058 T
  patches(4096)        4096 clusterlod-style patches of U[64,512]
                       triangles cut from a 2048x2048 grid
  request_stream(...)  serving stream: Poisson arrivals, log-uniform sizes
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

__all__ = [
    "Mesh",
    "grid",
    "icosphere",
    "seamed_grid",
    "patches",
    "request_stream",
    "target_count",
]


@dataclasses.dataclass
class Mesh:
    """Index buffer plus an interleaved float32 vertex buffer.

    `vertices` is [V, stride_floats]; positions are the first three floats, attributes (if
    any) follow at `attr_offset` floats with `attr_count` floats.
    """

    indices: np.ndarray  # uint32 [I]
    vertices: np.ndarray  # float32 [V, F]
    attr_offset: int = 3
    attr_count: int = 0
    name: str = ""

    @property
    def index_count(self) -> int:
        return int(self.indices.size)

    @property
    def vertex_count(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def triangle_count(self) -> int:
        return int(self.indices.size // 3)

    @property
    def stride(self) -> int:
        return int(self.vertices.shape[1] * 4)


def target_count(index_count: int, num: int, den: int) -> int:
    """target_index_count = (triangles * num / den) * 3, in integer arithmetic."""
    return (index_count // 3 * num // den) * 3


def _hash2(ix: np.ndarray, iy: np.ndarray, seed: int) -> np.ndarray:
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B1) + iy.astype(np.uint64) * np.uint64(0x85EBCA77)
         + np.uint64(seed & 0xFFFFFFFF) * np.uint64(0xC2B2AE3D)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    h = (h * np.uint64(0x2C1B3C6D)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(12)
    h = (h * np.uint64(0x297A2D39)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(15)
    return h


def value_noise(x: np.ndarray, y: np.ndarray, seed: int, freq: int = 16) -> np.ndarray:
    """Smoothstep-interpolated lattice value noise in [0, 1) (float64)."""
    fx = x * freq
    fy = y * freq
    ix = np.floor(fx).astype(np.int64)
    iy = np.floor(fy).astype(np.int64)
    tx = fx - ix
    ty = fy - iy
    sx = tx * tx * (3.0 - 2.0 * tx)
    sy = ty * ty * (3.0 - 2.0 * ty)

    def lat(a, b):
        return _hash2(a, b, seed).astype(np.float64) / 4294967296.0

    v00 = lat(ix, iy)
    v10 = lat(ix + 1, iy)
    v01 = lat(ix, iy + 1)
    v11 = lat(ix + 1, iy + 1)
    a = v00 + (v10 - v00) * sx
    b = v01 + (v11 - v01) * sx
    return a + (b - a) * sy


def _grid_indices(n: int) -> np.ndarray:
    """Two triangles per cell, (a, b, c) and (c, b, d), a=(i,j) b=(i+1,j) c=(i,j+1) d=(i+1,j+1)."""
    j, i = np.meshgrid(np.arange(n - 1, dtype=np.int64), np.arange(n - 1, dtype=np.int64), indexing="ij")
    a = j * n + i
    b = a + 1
    c = a + n
    d = c + 1
    tris = np.stack([a, b, c, c, b, d], axis=-1).reshape(-1)
    return tris.astype(np.uint32)


def _grid_positions(n: int, seed: int, amplitude: float = 0.1) -> np.ndarray:
    j, i = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    x = i / (n - 1)
    y = j / (n - 1)
    z = amplitude * value_noise(x, y, seed)
    return np.stack([x, y, z], axis=-1).reshape(-1, 3)


def grid(n: int = 256, seed: int = 1) -> Mesh:
    """Regular n x n grid in [0,1]^2 with a seeded value-noise height field of amplitude 0.1."""
    pos = _grid_positions(n, seed).astype(np.float32)
    return Mesh(_grid_indices(n), np.ascontiguousarray(pos), name=f"grid{n}")


def icosphere(levels: int = 8, seed: int = 2, amplitude: float = 0.05) -> Mesh:
    """Icosahedron subdivided `levels` times by normalised midpoints, radially displaced by
    r = 1 + amplitude * mean_k sin(f_k * dot(d_k, p) + phi_k), k = 1..8."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = np.array([
        [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
        [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
        [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)

    for _ in range(levels):
        nv = verts.shape[0]
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        edges = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        key = np.minimum(edges[:, 0], edges[:, 1]) * nv + np.maximum(edges[:, 0], edges[:, 1])
        uniq, inv = np.unique(key, return_inverse=True)
        e0 = uniq // nv
        e1 = uniq % nv
        mid = verts[e0] + verts[e1]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        verts = np.concatenate([verts, mid])
        m = inv.reshape(3, -1) + nv
        ab, bc, ca = m[0], m[1], m[2]
        faces = np.concatenate([
            np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])

    rng = np.random.Generator(np.random.PCG64(seed))
    d = rng.normal(size=(8, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    f = rng.integers(2, 16, size=8).astype(np.float64)
    phi = rng.uniform(0.0, 2.0 * math.pi, size=8)
    s = np.zeros(verts.shape[0])
    for k in range(8):
        s += np.sin(f[k] * (verts @ d[k]) + phi[k])
    r = 1.0 + amplitude * (s / 8.0)
    pos = (verts * r[:, None]).astype(np.float32)
    return Mesh(faces.reshape(-1).astype(np.uint32), np.ascontiguousarray(pos), name=f"icosphere{levels}")


def seamed_grid(n: int = 1024, seams: int = 16, seed: int = 3) -> Mesh:
    """n x n grid with per-vertex normal (central differences) and UV = (x, y); `seams` seeded
    grid lines (row or column in [1, n-2]) are cut into UV seams: the vertices on the line are
    duplicated, the cells on the far side reference the duplicates, and the duplicates' U (for
    a column) or V (for a row) is offset by 0.5.  Crossing lines give four-wedge vertices.
    Vertex layout: x y z nx ny nz u v (32-byte stride)."""
    pos = _grid_positions(n, seed)
    h = pos[:, 2].reshape(n, n)
    dzdx = np.empty_like(h)
    dzdy = np.empty_like(h)
    dzdx[:, 1:-1] = (h[:, 2:] - h[:, :-2]) * ((n - 1) / 2.0)
    dzdx[:, 0] = (h[:, 1] - h[:, 0]) * (n - 1)
    dzdx[:, -1] = (h[:, -1] - h[:, -2]) * (n - 1)
    dzdy[1:-1, :] = (h[2:, :] - h[:-2, :]) * ((n - 1) / 2.0)
    dzdy[0, :] = (h[1, :] - h[0, :]) * (n - 1)
    dzdy[-1, :] = (h[-1, :] - h[-2, :]) * (n - 1)
    nrm = np.stack([-dzdx.reshape(-1), -dzdy.reshape(-1), np.ones(n * n)], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    uv = pos[:, :2].copy()
    verts = np.concatenate([pos, nrm, uv], axis=1)

    idx = _grid_indices(n).astype(np.int64).reshape(-1, 3)
    # cell coordinates of every triangle (two per cell, row-major)
    cell = np.arange(idx.shape[0]) // 2
    cj = cell // (n - 1)
    ci = cell % (n - 1)
    vx = np.arange(n * n) % n  # column of a base vertex
    vy = np.arange(n * n) // n  # row of a base vertex

    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    lines = []
    for _ in range(seams):
        is_col = bool(rng.integers(0, 2))
        at = int(rng.integers(1, n - 1))
        lines.append((is_col, at))

    base_of = np.arange(n * n)  # base grid vertex of every vertex (duplicates inherit)
    for is_col, at in lines:
        cols = vx[base_of[idx]] if is_col else vy[base_of[idx]]
        side = (ci >= at) if is_col else (cj >= at)
        hit = (cols == at) & side[:, None]
        refs = np.unique(idx[hit])
        new_ids = np.arange(verts.shape[0], verts.shape[0] + refs.size)
        dup = verts[refs].copy()
        dup[:, 6 if is_col else 7] += 0.5
        verts = np.concatenate([verts, dup])
        base_of = np.concatenate([base_of, base_of[refs]])
        lut = np.full(verts.shape[0], -1, dtype=np.int64)
        lut[refs] = new_ids
        idx[hit] = lut[idx[hit]]

    return Mesh(idx.reshape(-1).astype(np.uint32), np.ascontiguousarray(verts.astype(np.float32)),
                attr_offset=3, attr_count=5, name=f"seamed_grid{n}")


SEAMED_GRID_WEIGHTS = np.array([0.5, 0.5, 0.5, 1.0, 1.0], dtype=np.float32)


def patches(count: int = 4096, n: int = 2048, tmin: int = 64, tmax: int = 512, seed: int = 4) -> list[Mesh]:
    """`count` clusterlod-style patches of U[tmin, tmax] triangles cut from a seeded displaced
    n x n grid: each patch is the first T triangles (row-major) of a random rectangle of cells,
    re-indexed locally in first-occurrence order with its own vertex buffer."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for p in range(count):
        t = int(rng.integers(tmin, tmax + 1))
        w = int(rng.integers(8, 33))
        hgt = (t + 2 * w - 1) // (2 * w)
        x0 = int(rng.integers(0, n - 1 - w))
        y0 = int(rng.integers(0, n - 1 - hgt))
        j, i = np.meshgrid(np.arange(y0, y0 + hgt), np.arange(x0, x0 + w), indexing="ij")
        a = (j * n + i).reshape(-1)
        b = a + 1
        c = a + n
        d = c + 1
        tri = np.stack([a, b, c, c, b, d], axis=-1).reshape(-1)[: t * 3]
        uniq, first = np.unique(tri, return_index=True)
        order = uniq[np.argsort(first, kind="stable")]
        lut = {int(v): k for k, v in enumerate(order)}
        local = np.fromiter((lut[int(v)] for v in tri), dtype=np.uint32, count=tri.size)
        gx = (order % n).astype(np.float64) / (n - 1)
        gy = (order // n).astype(np.float64) / (n - 1)
        gz = 0.1 * value_noise(gx, gy, seed, freq=64)
        pos = np.stack([gx, gy, gz], axis=-1).astype(np.float32)
        out.append(Mesh(local, np.ascontiguousarray(pos), name=f"patch{p}"))
    return out


@dataclasses.dataclass
class Request:
    arrival_s: float
    mesh: Mesh
    target_index_count: int
    weights: np.ndarray | None
    seed: int


def request_stream(count: int, rate_per_s: float, seed: int = 5, tmin: float = 1e3, tmax: float = 2e6) -> list[Request]:
    """Serving stream: Poisson arrivals at `rate_per_s`, sizes log-uniform in [tmin, tmax]
    triangles, half displaced grids and half icospheres (nearest subdivision level), target
    ratio from {0.5, 0.25, 0.1}, one third with normal+UV attributes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for r in range(count):
        t += float(rng.exponential(1.0 / rate_per_s))
        tris = math.exp(rng.uniform(math.log(tmin), math.log(tmax)))
        ratio = [(1, 2), (1, 4), (1, 10)][int(rng.integers(0, 3))]
        with_attr = int(rng.integers(0, 3)) == 0
        rseed = int(rng.integers(0, 1 << 30))
        if with_attr or rng.integers(0, 2) == 0:
            side = max(4, int(round(math.sqrt(tris / 2.0))) + 1)
            if with_attr:
                m = seamed_grid(side, seams=max(1, side // 64), seed=rseed)
            else:
                m = grid(side, seed=rseed)
        else:
            lv = max(1, min(9, int(round(math.log(tris / 20.0, 4)))))
            m = icosphere(lv, seed=rseed)
        w = SEAMED_GRID_WEIGHTS if m.attr_count else None
        out.append(Request(t, m, target_count(m.index_count, *ratio), w, rseed))
    return out
