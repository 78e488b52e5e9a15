"""Seeded synthetic meshes for the benchmark workloads.

Meshes are emitted the way a Gmsh MSH document lists them: node tags with
coordinates, and elements as lists of node tags in record order. The engine
(and the oracle) canonicalise them exactly as the reference's mesh builders
do (cpp/src/mesh/triangle/build_triangle_mesh.cpp, build_line_mesh.cpp).

Randomness is a counter-based splitmix64 stream, value k = mix(seed +
(k + 1) * golden), mapped to [0, 1) by (u64 >> 11) * 2^-53, so every mesh is
a pure function of (shape, seed) on any machine.
"""

from __future__ import annotations

import numpy as np

DEFAULT_SEED = 0x5EED
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Stream values offset .. offset+count-1 as uint64."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


class TriangleMeshInput:
    """Raw triangle mesh: node tags, node (x, y), triangles as node tags."""

    def __init__(self, node_tags, node_xy, triangles, box):
        self.node_tags = np.ascontiguousarray(node_tags, np.int64)
        self.node_xy = np.ascontiguousarray(node_xy, np.float64)
        self.triangles = np.ascontiguousarray(triangles, np.int64)
        self.box = box

    @property
    def element_count(self) -> int:
        return int(self.triangles.shape[0])


class LineMeshInput:
    """Raw line mesh: node tags, node x, segments as node tags."""

    def __init__(self, node_tags, node_x, segments):
        self.node_tags = np.ascontiguousarray(node_tags, np.int64)
        self.node_x = np.ascontiguousarray(node_x, np.float64)
        self.segments = np.ascontiguousarray(segments, np.int64)

    @property
    def element_count(self) -> int:
        return int(self.segments.shape[0])


def jittered_triangles(n: int, lower=(0.0, 0.0), upper=(1.0, 1.0), jitter: float = 0.15,
                       seed: int = DEFAULT_SEED) -> TriangleMeshInput:
    """2 n^2 triangles: an n x n quad grid on the box, one random diagonal per
    quad, interior nodes displaced uniformly by +/- jitter * h per axis.
    Boundary nodes stay on the grid so periodic edges match exactly.

    Stream layout: node k = j (n+1) + i draws values 2k (x) and 2k+1 (y);
    quad q = j n + i draws value 2 (n+1)^2 + q (< 0.5: diagonal ll-ur)."""
    nn = n + 1
    lx, ly = lower
    ux, uy = upper
    hx, hy = (ux - lx) / n, (uy - ly) / n
    j, i = np.meshgrid(np.arange(nn), np.arange(nn), indexing="ij")
    x = lx + (ux - lx) * i / n
    y = ly + (uy - ly) * j / n
    u = uniform(seed, 2 * nn * nn).reshape(nn, nn, 2)
    interior = (i > 0) & (i < n) & (j > 0) & (j < n)
    x = np.where(interior, x + (2.0 * u[..., 0] - 1.0) * jitter * hx, x)
    y = np.where(interior, y + (2.0 * u[..., 1] - 1.0) * jitter * hy, y)
    xy = np.stack([x.ravel(), y.ravel()], axis=1)
    tags = np.arange(1, nn * nn + 1, dtype=np.int64)

    d = uniform(seed, n * n, offset=2 * nn * nn).reshape(n, n) < 0.5
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ll = jj * nn + ii + 1
    lr = ll + 1
    ul = ll + nn
    ur = ul + 1
    t0 = np.where(d[..., None], np.stack([ll, lr, ur], -1), np.stack([ll, lr, ul], -1))
    t1 = np.where(d[..., None], np.stack([ll, ur, ul], -1), np.stack([lr, ur, ul], -1))
    tris = np.stack([t0, t1], axis=2).reshape(-1, 3)
    return TriangleMeshInput(tags, xy, tris, (lx, ly, ux, uy))


def jittered_line(cells: int, lower: float = 0.0, upper: float = 1.0, jitter: float = 0.20,
                  seed: int = DEFAULT_SEED) -> LineMeshInput:
    """cells segments on [lower, upper]; interior nodes displaced uniformly by
    +/- jitter * h (stream value k for node k)."""
    h = (upper - lower) / cells
    k = np.arange(cells + 1)
    x = lower + (upper - lower) * k / cells
    u = uniform(seed, cells + 1)
    interior = (k > 0) & (k < cells)
    x = np.where(interior, x + (2.0 * u - 1.0) * jitter * h, x)
    tags = np.arange(1, cells + 2, dtype=np.int64)
    seg = np.stack([tags[:-1], tags[1:]], axis=1)
    return LineMeshInput(tags, x, seg)
