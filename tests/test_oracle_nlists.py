"""Oracle neighbour lists (src/nlists/) against known answers.

* distance_matrix vs scipy KDTree on rect(20,20) (tests/test_neighbours.py:61-79);
* exclusion counts 44/38/26 for node 151 (tests/test_neighbours.py:30-41,81-92);
* point_distances pair sets vs brute-force O(N^2) enumeration with BFS
  exclusion sets, on a flattened sphere where repulsion pairs exist.
"""
import json
import os

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.spatial import KDTree

from oracle import oracle as O
from trimem_amd import meshes as M
import refmath

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


@pytest.fixture(scope="module")
def rect_mesh():
    v, f = M.rect(20, 20)
    return v, f, O.Mesh(v, f)


@pytest.mark.parametrize("ltype", [0, 1])
@pytest.mark.parametrize("rlist", GOLD["distance_matrix"]["rlist"])
def test_distance_matrix_vs_kdtree(rect_mesh, ltype, rlist):
    v, f, m = rect_mesh
    nl = O.NeighbourList(m, O.Params(rep_nlist=ltype, rep_rlist=rlist, rep_exclusion=0))
    d, i, j = nl.distance_matrix(m, GOLD["distance_matrix"]["dmax"])
    Mx = coo_matrix((d, (i, j)), shape=(len(v), len(v)))
    if ltype == 0:
        Mx = Mx + Mx.T
    tree = KDTree(v)
    C = tree.sparse_distance_matrix(tree, GOLD["distance_matrix"]["dmax"])
    assert np.allclose(C.todense(), Mx.todense())


@pytest.mark.parametrize("ltype", [0, 1])
@pytest.mark.parametrize("excl", [0, 1, 2])
def test_exclusion_counts(rect_mesh, ltype, excl):
    g = GOLD["exclusion_counts"]
    _, _, m = rect_mesh
    nl = O.NeighbourList(m, O.Params(rep_nlist=ltype, rep_rlist=g["rlist"], rep_exclusion=excl))
    _, jdx = nl.point_distances(m, g["pid"], g["rlist"])
    assert len(jdx) == g["counts"][str(excl)]


@pytest.mark.parametrize("ltype", [0, 1])
@pytest.mark.parametrize("excl", [0, 1, 2])
def test_point_distances_vs_brute_force(ltype, excl):
    v, f, lbar = M.jittered_icosphere(3, seed=5, rel_sigma=0.02)
    v = v * np.array([1.0, 0.8, 0.07])
    dmax = 1.5 * lbar
    m = O.Mesh(v, f)
    nl = O.NeighbourList(m, O.Params(rep_nlist=ltype, rep_rlist=dmax, rep_exclusion=excl))
    pairs = refmath.brute_pairs(v, f, dmax, excl)
    assert len(pairs) > 100
    want = {}
    for i, j in pairs:
        want.setdefault(int(i), set()).add(int(j))
    for pid in range(len(v)):
        d, jdx = nl.point_distances(m, pid, dmax)
        assert set(jdx.tolist()) == want.get(pid, set()), pid
        assert np.allclose(d, v[pid] - v[jdx], rtol=0, atol=0)


def test_cell_box_quirk():
    """cell_list.h:46-47 initialises only component 0 of the box bounds; for a
    mesh entirely at z > 0 the z-range is anchored at 0, which changes which
    vertices share a cell.  Pairs within dmax <= r_list are unaffected."""
    v, f, lbar = M.jittered_icosphere(2, seed=1)
    v = v + np.array([0.0, 5.0, 5.0])
    m = O.Mesh(v, f)
    nl = O.NeighbourList(m, O.Params(rep_nlist=0, rep_rlist=0.3, rep_exclusion=0))
    d, i, j = nl.distance_matrix(m, 0.3)
    tree = KDTree(v)
    C = tree.sparse_distance_matrix(tree, 0.3, output_type="ndarray")
    want = {(int(a), int(b)) for a, b, _ in C if a < b}
    assert {(int(a), int(b)) for a, b in zip(i, j)} == want
