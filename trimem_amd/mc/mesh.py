"""State wrapper around :class:`trimem_amd.core.TriMesh`
(src/trimem/mc/mesh.py:12-62).

``Mesh.x`` reads and writes the vertex positions through the mesh's
``points()`` view; ``Mesh.f`` returns the face-vertex indices.  After a
device-resident HMC step the view is refreshed from the GPU on access.
"""
import copy

import numpy as np

from .. import core as _core


class Mesh:
    """Vertices and triangles of one membrane (mesh.py:12-52).

    Keyword Args:
        points (ndarray[float]): (N,3) vertex positions.
        cells (ndarray[int]): (M,3) triangles.
    """

    def __init__(self, points=None, cells=None):
        if points is not None and cells is not None:
            self.trimesh = _core.TriMesh(points, cells)
        else:
            self.trimesh = _core.TriMesh()

    @property
    def x(self):
        """(N,3) vertex positions: a writable view owned by the mesh."""
        return self.trimesh.points()

    @x.setter
    def x(self, values):
        np.copyto(self.x, values)

    @property
    def f(self):
        """(M,3) face-vertex indices."""
        return self.trimesh.fv_indices()

    def copy(self):
        """Deep copy (like mesh.py:50-52, it needs a picklable TriMesh)."""
        return copy.deepcopy(self)


def read_trimesh(filename):
    """Mesh read from an STL, OFF or OBJ file (mesh.py:54-62)."""
    mesh = Mesh()
    mesh.trimesh = _core.read_mesh(filename)
    return mesh
