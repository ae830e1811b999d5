"""Oracle geometry primitives vs torch autograd on the facet pair.

Follows tests/test_properties.py:10-155 of the reference: a two-face patch
with dihedral angle +-pi/4, halfedge 0 (vertex 0 -> 1), tolerance 1e-10.
The torch functions below restate the definitions, not the oracle's
hand-written gradients, so the comparison pins mesh_util.h:15-185 as restated
in oracle/oracle.cpp.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O


def t_edge_length(p):
    return torch.linalg.norm(p[1] - p[0])


def t_face_normal_u(p):
    return torch.cross(p[0] - p[1], p[2] - p[1], dim=0)


def t_face_area(p):
    return torch.linalg.norm(t_face_normal_u(p)) / 2


def t_face_volume(p):
    return torch.dot(p[0], torch.cross(p[1], p[2], dim=0)) / 6


def t_dihedral(p):
    n0 = t_face_normal_u(p[[0, 1, 2]])
    n1 = t_face_normal_u(p[[1, 0, 3]])
    n0 = n0 / torch.linalg.norm(n0)
    n1 = n1 / torch.linalg.norm(n1)
    sgn = torch.dot(torch.cross(n0, n1, dim=0), p[1] - p[0])
    a = torch.acos(torch.dot(n0, n1))
    return a if sgn > 0 else -a


def value_and_grad(fn, pts):
    x = torch.tensor(pts, dtype=torch.float64, requires_grad=True)
    v = fn(x)
    (g,) = torch.autograd.grad(v, x)
    return float(v), g.numpy()


@pytest.fixture(params=[np.pi / 4, -np.pi / 4])
def facet_pair(request):
    alpha = request.param
    x = np.cos(alpha) * 0.5
    z = np.sin(alpha) * 0.5
    points = np.array([[0.0, -0.5, 0.0], [0.0, 0.5, 0.0], [-x, 0.0, z], [x, 0.0, z]])
    cells = np.array([[0, 1, 2], [1, 0, 3]])
    return points, cells, O.Mesh(points, cells)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def test_edge_length(facet_pair):
    p, _, m = facet_pair
    v, g = value_and_grad(t_edge_length, p[[0, 1]])
    assert abs(m.edge_length(0) - v) / v < 1e-10
    assert rel(m.edge_length_grad(0), g) < 1e-10


def test_face_area(facet_pair):
    p, _, m = facet_pair
    v, g = value_and_grad(t_face_area, p[[0, 1, 2]])
    assert abs(m.face_area(0) - v) / v < 1e-10
    assert rel(m.face_area_grad(0), g) < 1e-10


def test_face_volume(facet_pair):
    p, _, m = facet_pair
    v, g = value_and_grad(t_face_volume, p[[0, 1, 2]])
    assert abs(m.face_volume(0) - v) < 1e-10
    assert rel(m.face_volume_grad(0), g) < 1e-10


def test_dihedral_angle(facet_pair):
    p, _, m = facet_pair
    v, g = value_and_grad(t_dihedral, p)
    assert abs(m.dihedral_angle(0) - v) < 1e-10
    assert rel(m.dihedral_angle_grad(0), g) < 1e-10


def test_dihedral_magnitude_closed_form(facet_pair):
    """The wings sit at +-alpha from the plane of the shared edge's normal, so
    the face normals differ by 2|alpha|: |theta| = pi/2 for alpha = pi/4."""
    _, _, m = facet_pair
    assert abs(abs(m.dihedral_angle(0)) - np.pi / 2) < 1e-12


def test_halfedge_zero_is_first_face_edge(facet_pair):
    """OpenMesh add_face numbering: halfedge 0 runs 0 -> 1 (test_properties.py:94)."""
    _, _, m = facet_pair
    t = m.topology()
    assert t["he_to"][0] == 1 and t["he_to"][1] == 0
    assert t["he_face"][0] == 0 and t["he_face"][1] == 1
