"""Cell list far from the origin (cell_list.h:43-87).

The reference's box spans 0 in y and z (its bounds start at min = (+max, 0,
0), max = (-1, 0, 0)), so a mesh centred at (0, 200, 200) has a global grid
of more than 2^28 cells at an rlist of one or 1.5 edge lengths, of which the
mesh occupies a sliver; the reference keys occupied cells in a map.  The GPU table
covers only the occupied sub-grid, and these tests pin that it answers like
the full grid: pair sets and per-vertex candidate order bit-exact against
the oracle, energies and gradients within 1e-10, also for vertices moved out
of the sub-grid (between the mesh and the origin, and past the box).
"""
import numpy as np
import pytest

from oracle import oracle as O
from trimem_amd import meshes as M

from test_gpu_parity import base_params, make_pair, props_vec, rel, to_core_params

pytestmark = pytest.mark.gpu

OFFSET = np.array([0.0, 200.0, 200.0])


@pytest.fixture(scope="module")
def core():
    import trimem_amd.core as m
    assert m.device_count() > 0, "no GPU visible"
    return m


def reference_grid_cells(v, rlist, box_eps=1e-6):
    """Cell count of the reference's dense grid (cell_list.h:44-66)."""
    lo = np.array([v[:, 0].min(), min(0.0, v[:, 1].min()), min(0.0, v[:, 2].min())])
    hi = np.array([max(-1.0, v[:, 0].max()), max(0.0, v[:, 1].max()), max(0.0, v[:, 2].max())])
    n = 1
    for dim in hi - lo + box_eps:
        n *= max(int(dim / rlist), 1)
    return n


@pytest.fixture(scope="module")
def far_mesh():
    v, f, lbar = M.jittered_icosphere(5, seed=11, rel_sigma=0.02)
    # flattened so that the two caps repel across the gap
    v = v * np.array([1.0, 1.0, 0.02]) + OFFSET
    return v, f, lbar


@pytest.mark.parametrize("ltype", [0, 1])
def test_far_offset_pairs_bit_exact(core, far_mesh, ltype):
    v, f, lbar = far_mesh
    rl = lbar
    assert reference_grid_cells(v, rl) > 2**28
    p = O.Params(rep_nlist=ltype, rep_rlist=rl, rep_exclusion=1)
    om, gm = O.Mesh(v, f), core.TriMesh(v, f)
    onl = O.NeighbourList(om, p)
    gnl = core.make_nlist(gm, to_core_params(core, p))
    do, io, jo = onl.distance_matrix(om, rl)
    dg, ig, jg = gnl.distance_matrix(gm, rl)
    assert len(do) > len(v)
    assert sorted(zip(io.tolist(), jo.tolist(), do.tolist())) == sorted(zip(ig, jg, dg))
    for pid in range(0, len(v), 7):
        d0, j0 = onl.point_distances(om, pid, rl)
        d1, j1 = gnl.point_distances(gm, pid, rl)
        if ltype == 0:
            assert list(j0) == list(j1)
            d1 = np.array([[q.x, q.y, q.z] for q in d1]).reshape((-1, 3))
            assert np.array_equal(d0, d1)
        else:
            assert sorted(j0) == sorted(j1)


def test_far_offset_queries_leave_the_sub_grid(core, far_mesh):
    """Cells stay frozen at build time; vertices moved into the empty space
    between the mesh and the origin, or past the box, see empty stencils
    there, and the rest of the mesh still finds them where it looks."""
    v, f, lbar = far_mesh
    rl = 1.5 * lbar
    assert reference_grid_cells(v, rl) > 2**28
    p = O.Params(rep_nlist=0, rep_rlist=rl, rep_exclusion=0)
    om, gm = O.Mesh(v, f), core.TriMesh(v, f)
    onl = O.NeighbourList(om, p)
    gnl = core.make_nlist(gm, to_core_params(core, p))
    x = v.copy()
    moved = {10: [0.0, -100.0, 0.0],           # inside the global grid, outside the occupied cells
             20: [0.0, 0.0, -199.0],           # next to the origin
             30: [3.0 * rl + 2.0, 0.0, 0.0],   # past the box in x
             40: [0.0, -0.5 * rl, 0.0],        # within the sub-grid
             60: [0.0, 0.0, -400.0]}           # below the box in z
    for i, d in moved.items():
        x[i] += d
    # far beyond any cell: the reference's int conversion of the cell
    # coordinate overflows there, so only the GPU is asked (no neighbours)
    x[50, 0] -= 1e12
    gm.points()[:] = x
    om.x = x
    d1, j1 = gnl.point_distances(gm, 50, rl)
    assert len(j1) == 0 and len(d1) == 0
    for pid in list(moved) + [int(j) for j in onl.point_distances(om, 40, rl)[1]]:
        d0, j0 = onl.point_distances(om, pid, rl)
        d1, j1 = gnl.point_distances(gm, pid, rl)
        assert list(j0) == list(j1), pid
        d1 = np.array([[q.x, q.y, q.z] for q in d1]).reshape((-1, 3))
        assert np.array_equal(d0, d1), pid
    do, io, jo = onl.distance_matrix(om, rl)
    dg, ig, jg = gnl.distance_matrix(gm, rl)
    assert sorted(zip(io.tolist(), jo.tolist(), do.tolist())) == sorted(zip(ig, jg, dg))


@pytest.mark.parametrize("excl", [0, 2])
def test_far_offset_energy(core, far_mesh, excl):
    v, f, lbar = far_mesh
    lc1 = 1.5 * lbar
    pr = base_params(lbar, rep_lc1=lc1, rep_rlist=lc1, rep_nlist=0, rep_exclusion=excl, kappa_r=1.0,
                     kappa_a=1e2, kappa_v=1e2, kappa_t=1.0)
    assert reference_grid_cells(v, lc1) > 2**28
    om, oe, gm, ge = make_pair(core, v, f, pr)
    po, pg = oe.properties(om), props_vec(ge.properties(gm))
    assert po[5] > 0
    # the volume of a mesh at 200 from the origin sums face terms ~1e4 times
    # its value; the repulsion, bending and tether terms are translation free
    for k, tol in enumerate((1e-12, 1e-9, 1e-10, 1e-10, 1e-12, 1e-10)):
        assert abs(pg[k] - po[k]) <= tol * abs(po[k]), (k, pg[k], po[k])
    assert abs(ge.energy(gm) - oe.energy(om)) <= 1e-10 * abs(oe.energy(om))
    assert rel(ge.gradient(gm), oe.gradient(om)) <= 1e-10
    rng = np.random.default_rng(12)
    x = v + rng.normal(0, 0.2 * lbar, v.shape)
    gm.points()[:] = x
    om.x = x
    assert rel(ge.gradient(gm), oe.gradient(om)) <= 1e-10
    ge.update_repulsion(gm)
    oe.update_repulsion(om)
    assert abs(ge.energy(gm) - oe.energy(om)) <= 1e-10 * abs(oe.energy(om))
    assert rel(ge.gradient(gm), oe.gradient(om)) <= 1e-10
