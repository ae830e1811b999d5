"""Synthetic membrane meshes.

The reference takes its test and benchmark meshes from trimesh
(tests/util.py:71-79: ``trimesh.creation.icosphere`` / ``icosahedron``) and
generates tubes and rectangles itself (tests/util.py:5-69).  trimesh is not
installed here, so this module restates those generators:

* ``icosahedron()``: the 12-vertex icosahedron with trimesh's vertex and face
  tables, scaled to unit radius.
* ``icosphere(n)``: ``n`` rounds of midpoint subdivision of that icosahedron
  with trimesh's ordering of the new midpoint vertices (unique edges sorted by
  (max, min) vertex id) and its 4-way face split, then one projection onto the
  sphere, ``v += unit * (radius - |v|)``.  This is the construction of the
  trimesh release the reference documentation used: with it, the documented
  ``EnergyManager.energy`` value of doc/python-module.rst:88 is reproduced
  bit for bit (tests/test_oracle_energy.py).
* ``tube`` and ``rect``: the reference's own test generators.
* ``jittered_icosphere``: the benchmark input, vertices jittered by
  N(0, (0.01 * mean edge length)^2) from a seed.
"""
import numpy as np

__all__ = ["icosahedron", "icosphere", "tube", "rect", "jittered_icosphere", "mean_edge_length"]


def icosahedron():
    """Unit-radius icosahedron, vertex and face order as trimesh.creation.icosahedron."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0,
                  0, -1, t, 0, 1, t, 0, -1, -t, 0, 1, -t,
                  t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1], dtype=np.float64)
    f = np.array([0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11,
                  1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                  3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9,
                  4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1], dtype=np.int64)
    v = v.reshape((-1, 3)) / np.sqrt(2.0 + t)
    return v, f.reshape((-1, 3))


def _subdivide(vertices, faces):
    """One round of 4-way midpoint subdivision (trimesh.remesh.subdivide order)."""
    edges = faces[:, [0, 1, 1, 2, 2, 0]].reshape((-1, 2))
    edges = np.sort(edges, axis=1)
    # trimesh hashes sorted edge rows as a ^ (b << 32) and orders unique rows
    # by that hash, i.e. by (max, min)
    h = edges[:, 0].astype(np.int64) ^ (edges[:, 1].astype(np.int64) << 32)
    _, unique, inverse = np.unique(h, return_index=True, return_inverse=True)
    mid = vertices[edges[unique]].mean(axis=1)
    mid_idx = inverse.reshape((-1, 3)) + len(vertices)
    f = np.column_stack([faces[:, 0], mid_idx[:, 0], mid_idx[:, 2],
                         mid_idx[:, 0], faces[:, 1], mid_idx[:, 1],
                         mid_idx[:, 2], mid_idx[:, 1], faces[:, 2],
                         mid_idx[:, 0], mid_idx[:, 1], mid_idx[:, 2]]).reshape((-1, 3))
    return np.vstack((vertices, mid)), f


def icosphere(subdivisions=3, radius=1.0):
    """Icosphere built like trimesh.creation.icosphere."""
    v, f = icosahedron()
    for _ in range(subdivisions):
        v, f = _subdivide(v, f)
    scalar = np.sqrt((v ** 2).sum(axis=1))
    unit = v / scalar.reshape((-1, 1))
    offset = radius - scalar
    v = v + unit * offset.reshape((-1, 1))
    return v, f


def tube(length=1.0, radius=1.0, segments=30):
    """Open tube around the y-axis (tests/util.py:5-43)."""
    U = 2 * np.pi * radius
    t = U / segments
    n = int(length // t) + 1
    angles = np.linspace(0, 2 * np.pi, segments + 1)[:-1]
    y = np.linspace(-length / 2, length / 2, n)
    x = np.cos(angles) * radius
    z = np.sin(angles) * radius
    verts = np.hstack((np.tile(x, (1, n)).T,
                       np.repeat(y, segments).reshape((segments * n, 1)),
                       np.tile(z, (1, n)).T))
    idx = np.arange(segments)
    faces = np.vstack((np.array([idx, idx + segments, (idx + 1) % segments]).T,
                       np.array([(idx + 1) % segments, idx + segments,
                                 (idx + 1) % segments + segments]).T))
    faces = np.vstack([faces + (i * segments) for i in range(n - 1)])
    return verts, faces


def rect(nx, ny):
    """Flat rectangular patch in the z=0 plane (tests/util.py:46-69)."""
    xx, yy = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny))
    verts = np.hstack((xx.reshape((nx * ny, 1)), yy.reshape((nx * ny, 1)),
                       np.zeros((nx * ny, 1))))
    ii = np.arange(nx)
    faces = (np.array([ii[:nx - 1], ii[1:nx], ii[1:nx] + nx]).T,
             np.array([ii[:nx - 1] + nx, ii[:nx - 1], ii[1:nx] + nx]).T)
    faces = np.vstack(faces)
    faces = np.vstack([faces + (i * nx) for i in range(ny - 1)])
    return verts, faces


def mean_edge_length(v, f):
    """Mean length over unique edges (same value as trimem's avg_tri_props l)."""
    e = np.sort(f[:, [0, 1, 1, 2, 2, 0]].reshape((-1, 2)), axis=1)
    e = np.unique(e, axis=0)
    return float(np.linalg.norm(v[e[:, 0]] - v[e[:, 1]], axis=1).mean())


def jittered_icosphere(level, seed, rel_sigma=0.01):
    """Unit icosphere with vertices jittered by N(0, (rel_sigma * mean edge)^2)."""
    v, f = icosphere(level)
    lbar = mean_edge_length(v, f)
    rng = np.random.default_rng(seed)
    v = v + rng.normal(0.0, rel_sigma * lbar, size=v.shape)
    return np.ascontiguousarray(v), np.ascontiguousarray(f.astype(np.int32)), lbar
