"""CPU checks of the C ABI library: it loads and exports every function that
include/trimem_hip.h declares; host-side mesh construction matches the
oracle's OpenMesh numbering (no GPU calls)."""
import ctypes

import numpy as np

from oracle import oracle as O
from trimem_amd import capi, meshes


def test_exports_every_declared_symbol():
    L = ctypes.CDLL(capi.LIB_PATH)
    names = capi.declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_core_module_imports_and_mesh_topology():
    import trimem_amd.core as m
    for v, f in (meshes.icosphere(2), meshes.tube(1.0, 0.5, 12), meshes.rect(6, 5)):
        t = m.TriMesh(v, f)
        o = O.Mesh(v, f)
        assert t.n_vertices() == o.n_vertices() and t.n_edges() == o.n_edges()
        assert np.array_equal(t.fv_indices(), o.fv_indices())
        for h in range(0, 2 * t.n_edges(), 5):
            assert t.next_halfedge_handle(t.halfedge_handle(h)).idx() == o.next_halfedge(h)


def test_points_view_is_writable_and_shared():
    import trimem_amd.core as m
    v, f = meshes.icosphere(1)
    t = m.TriMesh(v, f)
    x = t.points()
    assert x.shape == v.shape and np.array_equal(x, v)
    x[3, 1] = 7.5
    assert t.points()[3, 1] == 7.5


def test_shape_errors():
    import pytest
    import trimem_amd.core as m
    with pytest.raises(RuntimeError):
        m.TriMesh(np.zeros((4, 2)), np.zeros((1, 3), dtype=np.int32))
    with pytest.raises(RuntimeError):
        m.TriMesh(np.zeros((4, 3)), np.zeros((1, 2), dtype=np.int32))
    # out-of-range ids are dropped, as in from_points_cells
    t = m.TriMesh(np.eye(3), np.array([[0, 1, 2], [0, 1, 9]], dtype=np.int32))
    assert t.fv_indices().shape == (1, 3)
