"""Meshes the reference's own tests build in code, restated as raw records.

These are fixtures: closed-form node coordinates and element records from
the reference tests, reproduced with Python's libm so values match the C++
std::sin/std::cos the reference uses.
"""

from __future__ import annotations

import math

import numpy as np


def irregular_line(count: int):
    """cpp/tests/runtime/case_regression_test.cpp:77-95: nodes perturbed off a
    uniform grid, elements listed in reverse with alternating orientation."""
    tags, xs = [], []
    for node in range(count + 1):
        x = node / count
        if 0 < node < count:
            x += 0.3 / count * math.sin(7.0 * node)
        tags.append(node + 1)
        xs.append(x)
    segs = []
    for element in range(count - 1, -1, -1):
        first, second = element + 1, element + 2
        segs.append([second, first] if element % 2 else [first, second])
    return np.array(tags, np.int64), np.array(xs), np.array(segs, np.int64)


def irregular_box(n: int):
    """cpp/tests/runtime/case_regression_test.cpp:98-129: unit box, perturbed
    interior nodes, alternating diagonals."""
    def ident(i, j):
        return j * (n + 1) + i + 1

    tags, xy = [], []
    for j in range(n + 1):
        for i in range(n + 1):
            x, y = i / n, j / n
            if 0 < i < n and 0 < j < n:
                x += 0.2 / n * math.sin(3.0 * i + 5.0 * j)
                y += 0.2 / n * math.cos(2.0 * i - 7.0 * j)
            tags.append(ident(i, j))
            xy.append((x, y))
    tris = []
    for j in range(n):
        for i in range(n):
            if (i + j) % 2 == 0:
                tris.append([ident(i, j), ident(i + 1, j), ident(i + 1, j + 1)])
                tris.append([ident(i, j), ident(i, j + 1), ident(i + 1, j + 1)])
            else:
                tris.append([ident(i, j), ident(i + 1, j), ident(i, j + 1)])
                tris.append([ident(i + 1, j), ident(i + 1, j + 1), ident(i, j + 1)])
    return np.array(tags, np.int64), np.array(xy), np.array(tris, np.int64)


def unit_square(cells: int):
    """cpp/tests/dg/euler_triangle_residual_test.cpp:94-120: structured unit
    square, two triangles per cell along the lower-left/upper-right diagonal."""
    def node(i, j):
        return j * (cells + 1) + i + 1

    tags, xy = [], []
    for j in range(cells + 1):
        for i in range(cells + 1):
            tags.append(node(i, j))
            xy.append((i / cells, j / cells))
    tris = []
    for j in range(cells):
        for i in range(cells):
            tris.append([node(i, j), node(i + 1, j), node(i + 1, j + 1)])
            tris.append([node(i, j), node(i + 1, j + 1), node(i, j + 1)])
    return np.array(tags, np.int64), np.array(xy), np.array(tris, np.int64)


def read_case_regression(path):
    """Parse cpp/tests/runtime/case_regression_reference.txt rows."""
    rows = {}
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            name, step, time, csum, cl2, mass = line.split()
            rows[name] = (int(step), float(time), float(csum), float(cl2), float(mass))
    return rows
