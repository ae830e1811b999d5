"""GPU parity: trimem_amd.core (HIP kernels via the C ABI) vs the FP64 oracle.

Tolerances (stated per check):
  * FP64 path: global properties and energy rel 1e-10, gradient
    ||g - g_ref|| / ||g_ref|| <= 1e-10 (the two differ only in summation
    order and in the regrouped gradient formula);
  * mixed path: gradient rel <= 1e-6, energy rel <= 1e-9;
  * neighbour lists: pair sets and distances bit-exact;
  * serial flips with injected streams: topology (fv_indices) bit-exact,
    same accepted and consumed-uniform counts.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from trimem_amd import meshes as M

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


@pytest.fixture(scope="module")
def core():
    import trimem_amd.core as m
    assert m.device_count() > 0, "no GPU visible"
    return m


def to_core_params(m, p):
    e = m.EnergyParams()
    for k in ("kappa_b", "kappa_a", "kappa_v", "kappa_c", "kappa_t", "kappa_r",
              "area_frac", "volume_frac", "curvature_frac"):
        setattr(e, k, getattr(p, k))
    b = m.BondParams()
    b.lc0, b.lc1, b.a0, b.r = p.bond_lc0, p.bond_lc1, p.bond_a0, p.bond_r
    b.type = [m.BondType.Edge, m.BondType.Area, getattr(m.BondType, "None")][p.bond_type]
    e.bond_params = b
    r = m.SurfaceRepulsionParams()
    r.lc1, r.r, r.rlist, r.exclusion_level = p.rep_lc1, p.rep_r, p.rep_rlist, p.rep_exclusion
    r.n_search = ["cell-list", "verlet-list"][p.rep_nlist]
    e.repulse_params = r
    c = m.ContinuationParams()
    c.delta, c.lam = p.cont_delta, p.cont_lambda
    e.continuation_params = c
    return e


def props_vec(vp):
    return np.array([vp.area, vp.volume, vp.curvature, vp.bending, vp.tethering, vp.repulsion])


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def base_params(lbar, **kw):
    d = dict(kappa_b=30.0, kappa_a=1e6, kappa_v=1e6, kappa_c=1e3, kappa_t=1e5, kappa_r=1e3,
             area_frac=0.95, volume_frac=0.8, curvature_frac=1.05, bond_type=0,
             bond_lc0=1.02 * lbar, bond_lc1=0.98 * lbar, bond_a0=1.0, bond_r=2)
    d.update(kw)
    return O.Params(**d)


def make_pair(core, v, f, p):
    om = O.Mesh(v, f)
    oe = O.EnergyManager(om, p)
    gm = core.TriMesh(v, f)
    ge = core.EnergyManager(gm, to_core_params(core, p))
    return om, oe, gm, ge


def check_fp64(core, v, f, p, tol=1e-10):
    om, oe, gm, ge = make_pair(core, v, f, p)
    assert rel(props_vec(ge.initial_props), oe.initial_props()) < tol
    po = oe.properties(om)
    pg = props_vec(ge.properties(gm))
    for k in range(6):
        assert abs(pg[k] - po[k]) <= tol * max(abs(po[k]), 1e-12), (k, pg[k], po[k])
    eo, eg = oe.energy(om), ge.energy(gm)
    assert abs(eg - eo) <= tol * abs(eo)
    go, gg = oe.gradient(om), ge.gradient(gm)
    assert rel(gg, go) <= tol, rel(gg, go)
    return om, oe, gm, ge


# ---------------------------------------------------------- primitives -----
@pytest.mark.parametrize("alpha", [np.pi / 4, -np.pi / 4])
def test_halfedge_primitives(core, alpha):
    x, z = np.cos(alpha) * 0.5, np.sin(alpha) * 0.5
    pts = np.array([[0.0, -0.5, 0.0], [0.0, 0.5, 0.0], [-x, 0.0, z], [x, 0.0, z]])
    cells = np.array([[0, 1, 2], [1, 0, 3]])
    om, gm = O.Mesh(pts, cells), core.TriMesh(pts, cells)
    for h in range(2 * gm.n_edges()):
        he = gm.halfedge_handle(h)
        assert abs(core.edge_length(gm, he) - om.edge_length(h)) < 1e-14
        assert rel(core.edge_length_grad(gm, he), om.edge_length_grad(h)) < 1e-14
        if gm.n_edges() and h in (0, 1):
            assert abs(core.face_area(gm, he) - om.face_area(h)) < 1e-14
            assert rel(core.face_area_grad(gm, he), om.face_area_grad(h)) < 1e-13
            assert abs(core.face_volume(gm, he) - om.face_volume(h)) < 1e-14
            assert rel(core.face_volume_grad(gm, he), om.face_volume_grad(h)) < 1e-13
        assert abs(core.dihedral_angle(gm, he) - om.dihedral_angle(h)) < 1e-13
    he = gm.halfedge_handle(0)
    assert rel(core.dihedral_angle_grad(gm, he), om.dihedral_angle_grad(0)) < 1e-12


# ------------------------------------------------------- energy, FP64 ------
@pytest.mark.parametrize("bond", [0, 1])
def test_fp64_icosphere(core, bond):
    v, f, lbar = M.jittered_icosphere(3, seed=20261202, rel_sigma=0.05)
    a = 4 * np.pi / len(f)
    check_fp64(core, v, f, base_params(lbar, bond_type=bond, bond_a0=a))


@pytest.mark.parametrize("nlist,excl", [(0, 0), (0, 1), (0, 2), (1, 0), (1, 2)])
def test_fp64_repulsion(core, nlist, excl):
    v, f, lbar = M.jittered_icosphere(3, seed=3, rel_sigma=0.02)
    v = v * np.array([1.0, 1.0, 0.06])
    lc1 = 2.0 * lbar
    p = base_params(lbar, rep_lc1=lc1, rep_rlist=lc1, rep_nlist=nlist, rep_exclusion=excl, kappa_r=1.0,
                    kappa_a=1e2, kappa_v=1e2, kappa_t=1.0)
    om, oe, gm, ge = check_fp64(core, v, f, p)
    assert oe.properties(om)[5] > 0


def test_fp64_open_tube(core):
    """Boundary vertices take the OpenMesh-semantics walker (reference
    boundary branches included)."""
    v, f = M.tube(length=1.0, radius=0.5, segments=24)
    lbar = M.mean_edge_length(v, f)
    rng = np.random.default_rng(4)
    v = v + rng.normal(0, 0.02 * lbar, v.shape)
    check_fp64(core, v, f, base_params(lbar, kappa_a=10.0, kappa_v=10.0, kappa_t=1.0, volume_frac=0.9))


def test_doc_energy_value(core):
    v, f = M.icosphere(3)
    v = v.astype(np.float32).astype(np.float64)
    gm = core.TriMesh(v, f)
    p = O.Params(**GOLD["doc_energy"]["params"])
    ge = core.EnergyManager(gm, to_core_params(core, p))
    ref = GOLD["doc_energy"]["energy"]
    assert abs(ge.energy(gm) - ref) / ref < 1e-12
    g = ge.gradient(gm)
    for vid, r in zip((0, 532, 196), GOLD["doc_energy"]["gradient_rows_first"]):
        assert np.linalg.norm(g[vid] - r) / np.linalg.norm(r) < 5e-9


def test_points_view_and_continuation(core):
    v, f, lbar = M.jittered_icosphere(3, seed=8, rel_sigma=0.05)
    p = base_params(lbar, cont_delta=0.25, cont_lambda=0.0)
    om, oe, gm, ge = make_pair(core, v, f, p)
    x = gm.points()
    x += 0.01 * np.sin(7 * x)
    om.x = x
    for _ in range(3):
        ge.update_reference_properties()
        oe.update_reference_properties()
    assert abs(ge.eparams.continuation_params.lam - 0.75) < 1e-15
    assert abs(ge.energy(gm) - oe.energy(om)) <= 1e-10 * abs(oe.energy(om))
    assert rel(ge.gradient(gm), oe.gradient(om)) <= 1e-10


def test_gradient_deterministic(core):
    v, f, lbar = M.jittered_icosphere(4, seed=9)
    gm = core.TriMesh(v, f)
    ge = core.EnergyManager(gm, to_core_params(core, base_params(lbar)))
    g1, g2 = ge.gradient(gm), ge.gradient(gm)
    assert np.array_equal(g1, g2)


# ------------------------------------------------------------ mixed --------
@pytest.mark.parametrize("level", [4, 5])
def test_mixed_precision(core, level):
    v, f, lbar = M.jittered_icosphere(level, seed=20261203)
    p = base_params(lbar, kappa_c=0.0, area_frac=1.0, volume_frac=0.8, curvature_frac=1.0,
                    bond_lc0=1.25 * lbar, bond_lc1=0.75 * lbar, rep_lc1=1.5 * lbar, rep_rlist=1.5 * lbar)
    om, oe, gm, ge = make_pair(core, v, f, p)
    ge.precision = "mixed"
    go = oe.gradient(om)
    gg = ge.gradient(gm)
    err = rel(gg, go)
    print(f"mixed gradient rel err L{level}: {err:.3e}")
    assert err <= 1e-6
    eo = oe.energy(om)
    assert abs(ge.energy(gm) - eo) <= 1e-9 * abs(eo)


# ----------------------------------------------------- neighbour lists -----
@pytest.mark.parametrize("ltype", [0, 1])
@pytest.mark.parametrize("excl", [0, 1, 2])
def test_exclusion_counts(core, ltype, excl):
    g = GOLD["exclusion_counts"]
    v, f = M.rect(20, 20)
    gm = core.TriMesh(v, f)
    p = O.Params(rep_nlist=ltype, rep_rlist=g["rlist"], rep_exclusion=excl)
    nl = core.make_nlist(gm, to_core_params(core, p))
    _, jdx = nl.point_distances(gm, g["pid"], g["rlist"])
    assert len(jdx) == g["counts"][str(excl)]


@pytest.mark.parametrize("ltype", [0, 1])
@pytest.mark.parametrize("excl", [0, 2])
def test_nlist_bit_exact(core, ltype, excl):
    v, f, lbar = M.jittered_icosphere(3, seed=5, rel_sigma=0.02)
    v = v * np.array([1.0, 0.8, 0.07])
    rl = 1.5 * lbar
    p = O.Params(rep_nlist=ltype, rep_rlist=rl, rep_exclusion=excl)
    om, gm = O.Mesh(v, f), core.TriMesh(v, f)
    onl = O.NeighbourList(om, p)
    gnl = core.make_nlist(gm, to_core_params(core, p))
    for dmax in (rl, 0.7 * rl):
        do, io, jo = onl.distance_matrix(om, dmax)
        dg, ig, jg = gnl.distance_matrix(gm, dmax)
        a = sorted(zip(io.tolist(), jo.tolist(), do.tolist()))
        b = sorted(zip(ig, jg, dg))
        assert a == b
    for pid in range(0, len(v), 3):
        d0, j0 = onl.point_distances(om, pid, rl)
        d1, j1 = gnl.point_distances(gm, pid, rl)
        assert list(j0) == list(j1) if ltype == 0 else sorted(j0) == sorted(j1)
        d1 = np.array([[q.x, q.y, q.z] for q in d1]).reshape((-1, 3))
        if ltype == 0:
            assert np.array_equal(d0, d1)


def test_nlist_after_motion(core):
    """Cells stay frozen between rebuilds; queries use current positions."""
    v, f, lbar = M.jittered_icosphere(3, seed=6, rel_sigma=0.02)
    v = v * np.array([1.0, 1.0, 0.08])
    rl = 2.0 * lbar
    p = base_params(lbar, rep_lc1=rl, rep_rlist=rl, rep_nlist=0, rep_exclusion=2, kappa_r=1.0,
                    kappa_a=1e2, kappa_v=1e2, kappa_t=1.0)
    om, oe, gm, ge = make_pair(core, v, f, p)
    rng = np.random.default_rng(1)
    x = v + rng.normal(0, 0.3 * lbar, v.shape)
    gm.points()[:] = x
    om.x = x
    assert abs(ge.energy(gm) - oe.energy(om)) <= 1e-10 * abs(oe.energy(om))
    assert rel(ge.gradient(gm), oe.gradient(om)) <= 1e-10
    ge.update_repulsion(gm)
    oe.update_repulsion(om)
    assert rel(ge.gradient(gm), oe.gradient(om)) <= 1e-10


# ---------------------------------------------------------------- utils ----
def test_mesh_utils(core):
    v, f, lbar = M.jittered_icosphere(4, seed=2)
    om, gm = O.Mesh(v, f), core.TriMesh(v, f)
    assert abs(core.area(gm) - om.area()) < 1e-12 * om.area()
    assert abs(core.edges_length(gm) - om.edges_length()) < 1e-12 * om.edges_length()
    a, l = core.avg_tri_props(gm)
    ao, lo = om.avg_tri_props()
    assert abs(a - ao) < 1e-12 * ao and abs(l - lo) < 1e-12 * lo


# ---------------------------------------------------------------- flips ----
def flip_case(level, seed, rep, nlist=0, excl=1):
    v, f, lbar = M.jittered_icosphere(level, seed=seed, rel_sigma=0.05)
    kw = dict(kappa_b=1.0, kappa_a=1.0, kappa_v=1.0, kappa_c=1.0, kappa_t=1.0, kappa_r=1.0,
              area_frac=1.0, volume_frac=1.0, curvature_frac=1.0, bond_lc0=1.15 * lbar, bond_lc1=0.85 * lbar)
    if rep:
        kw.update(rep_lc1=1.5 * lbar, rep_rlist=1.5 * lbar, rep_exclusion=excl, rep_nlist=nlist)
    return v, f, O.Params(**kw)


@pytest.mark.parametrize("level,rep,nlist,excl", [(2, False, 0, 1), (3, False, 0, 1), (3, True, 0, 1),
                                                  (3, True, 0, 2), (3, True, 0, 0), (3, True, 1, 0), (3, True, 1, 1),
                                                  (5, True, 0, 2)])
def test_flip_serial_injected_parity(core, level, rep, nlist, excl):
    """Accepted count, uniforms used and the face list after a serial sweep
    over an injected proposal/uniform stream equal the oracle's, for the
    cell list at exclusion levels 0-2 and the Verlet list."""
    v, f, p = flip_case(level, 11 + level + 3 * (excl - 1) + 7 * nlist, rep, nlist, excl)
    om, oe, gm, ge = make_pair(core, v, f, p)
    rng = np.random.default_rng(level)
    n = int(gm.n_edges() * 0.5)
    prop = rng.integers(0, gm.n_edges(), n).astype(np.int32)
    uni = rng.random(n)
    acc_o, used_o = oe.flip_serial_injected(om, prop, uni)
    acc_g, used_g = core.flip_injected(gm, ge, prop, uni)
    assert (acc_g, used_g) == (acc_o, used_o)
    assert np.array_equal(gm.fv_indices(), om.fv_indices())
    # energy and gradient agree on the flipped topology
    assert abs(ge.energy(gm) - oe.energy(om)) <= 1e-10 * abs(oe.energy(om))
    assert rel(ge.gradient(gm), oe.gradient(om)) <= 1e-10


def with_window(core, gm, p, w):
    ge = core.EnergyManager(gm, to_core_params(core, p))
    t = ge.tuning
    t["serial_window"] = w
    ge.tuning = t
    assert ge.tuning["serial_window"] == w
    return ge


@pytest.mark.parametrize("rep,excl", [(True, 2), (False, 1)])
def test_serial_flip_windows_match_one_workgroup(core, rep, excl):
    """The serial sweep in windows of W proposals (evaluated at once, then
    committed in order up to the first proposal that an earlier flip of the
    window could have changed) takes the one-workgroup sweep's decisions and
    topology bit for bit.  Level 5 with ~73% acceptance, at exclusion level 2
    (the neighbour-set rule) and without repulsion."""
    v, f, p = flip_case(5, 26, rep, 0, excl)
    runs = {}
    for w in (0, 1, 7, 64, 256):
        gm = core.TriMesh(v, f)
        ge = with_window(core, gm, p, w)
        accs = [core.flip(gm, ge, 0.3, seed=99 + k) for k in range(2)]
        runs[w] = (accs, gm.fv_indices().copy(), ge.energy(gm))
        print(w, accs, flush=True)
    assert min(runs[0][0]) > 0.1 * len(f)
    for w in (1, 7, 64, 256):
        assert runs[w][0] == runs[0][0], w
        assert np.array_equal(runs[w][1], runs[0][1]), w
        assert runs[w][2] == runs[0][2], w


@pytest.mark.parametrize("w", [0, 7, 256])
def test_flip_serial_open_tube(core, w):
    """Serial sweeps over an injected stream on an open tube (boundary edges
    are never flipped, and a boundary proposal's window entry covers only
    its endpoints) give the oracle's accepted count, consumed uniforms and
    faces, for the one-workgroup sweep and windows of 7 and 256; cell list
    at exclusion level 2."""
    v, f = M.tube(length=1.0, radius=0.5, segments=24)
    lbar = M.mean_edge_length(v, f)
    v = v + np.random.default_rng(3).normal(0, 0.03 * lbar, v.shape)
    p = O.Params(kappa_b=1.0, kappa_a=1.0, kappa_v=1.0, kappa_c=1.0, kappa_t=1.0, kappa_r=1.0,
                 bond_lc0=1.25 * lbar, bond_lc1=0.75 * lbar, rep_lc1=1.5 * lbar, rep_rlist=1.5 * lbar,
                 rep_exclusion=2)
    om = O.Mesh(v, f)
    oe = O.EnergyManager(om, p)
    gm = core.TriMesh(v, f)
    ge = with_window(core, gm, p, w)
    ne = gm.n_edges()
    rng = np.random.default_rng(4)
    prop = rng.integers(0, ne, ne).astype(np.int32)
    uni = rng.random(ne)
    acc_o, used_o = oe.flip_serial_injected(om, prop, uni)
    acc_g, used_g = core.flip_injected(gm, ge, prop, uni)
    print(w, acc_o, used_o, flush=True)
    assert used_o < ne and acc_o > 0          # boundary proposals consume no uniform
    assert (acc_g, used_g) == (acc_o, used_o)
    assert np.array_equal(gm.fv_indices(), om.fv_indices())


def test_pflip_open_tube_consistent_with_oracle(core):
    """Parallel sweeps on an open tube (boundary edges are never claimed)
    leave a device topology whose properties, energy and gradient match an
    oracle mesh rebuilt from its faces (area/volume/curvature penalties off,
    so the two managers' reference values do not enter)."""
    v, f = M.tube(length=1.0, radius=0.5, segments=24)
    lbar = M.mean_edge_length(v, f)
    v = v + np.random.default_rng(5).normal(0, 0.03 * lbar, v.shape)
    p = O.Params(kappa_b=1.0, kappa_a=0.0, kappa_v=0.0, kappa_c=0.0, kappa_t=1.0, kappa_r=1.0,
                 bond_lc0=1.25 * lbar, bond_lc1=0.75 * lbar, rep_lc1=1.5 * lbar, rep_rlist=1.5 * lbar,
                 rep_exclusion=2)
    gm = core.TriMesh(v, f)
    ge = core.EnergyManager(gm, to_core_params(core, p))
    accs = [core.pflip(gm, ge, 0.5, seed=300 + k) for k in range(4)]
    fv = gm.fv_indices()
    assert sum(accs) > 0 and not np.array_equal(fv, f)
    om = O.Mesh(v, fv)
    oe = O.EnergyManager(om, p)
    assert rel(props_vec(ge.properties(gm)), oe.properties(om)) < 1e-10
    assert abs(ge.energy(gm) - oe.energy(om)) <= 1e-10 * abs(oe.energy(om))
    assert rel(ge.gradient(gm), oe.gradient(om)) < 1e-10


def test_serial_flip_errors_match_one_workgroup(core):
    """A proposal index out of range stops the sweep at that proposal, and an
    exhausted uniform stream after the proposal that found it empty (which
    counts as rejected): the windowed sweep raises the same error as the
    one-workgroup sweep and leaves the same topology."""
    v, f, p = flip_case(4, 27, False)
    ne = len(f) * 3 // 2
    rng = np.random.default_rng(5)
    prop = rng.integers(0, ne, 2000).astype(np.int32)
    bad = prop.copy()
    bad[1500] = ne + 3
    uni = rng.random(2000)
    for args in ((bad, uni), (prop, uni[:700])):
        outs = []
        for w in (0, 256):
            gm = core.TriMesh(v, f)
            ge = with_window(core, gm, p, w)
            with pytest.raises(RuntimeError) as ei:
                core.flip_injected(gm, ge, *args)
            outs.append((str(ei.value), gm.fv_indices().copy(), ge.energy(gm)))
        print(outs[0][0], flush=True)
        assert outs[0][0] == outs[1][0]
        assert np.array_equal(outs[0][1], outs[1][1])
        assert outs[0][2] == outs[1][2]
        assert not np.array_equal(outs[0][1], f)


def test_flip_smoke_reference_bound(core):
    """tests/test_flips.py:52-60: > 100 accepted at flip_ratio 0.5."""
    v, f = M.icosphere(3)
    gm = core.TriMesh(v, f)
    a, l = core.avg_tri_props(gm)
    p = O.Params(kappa_b=1.0, kappa_a=1.0, kappa_v=1.0, kappa_c=1.0, kappa_t=1.0,
                 bond_lc0=1.15 * l, bond_lc1=0.85 * l, bond_a0=a)
    ge = core.EnergyManager(gm, to_core_params(core, p))
    assert core.flip(gm, ge, 0.5, seed=1) > 100


# ------------------------------------------------------ parallel flips -----
def check_manifold(f, nv, ne):
    assert nv - ne + len(f) == 2
    dirs = set()
    for a, b, c in f:
        assert len({a, b, c}) == 3
        for x, y in ((a, b), (b, c), (c, a)):
            assert (x, y) not in dirs
            dirs.add((x, y))
    for (x, y) in dirs:
        assert (y, x) in dirs


def test_pflip_deterministic_and_manifold(core):
    v, f, p = flip_case(3, 21, True)
    runs = []
    for _ in range(2):
        gm = core.TriMesh(v, f)
        ge = core.EnergyManager(gm, to_core_params(core, p))
        accs = [core.pflip(gm, ge, 0.3, seed=1234 + k) for k in range(5)]
        runs.append((accs, gm.fv_indices().copy()))
        check_manifold(gm.fv_indices(), gm.n_vertices(), gm.n_edges())
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])
    assert sum(runs[0][0]) > 0


def test_pflip_scan_window_takes_sequential_decisions(core):
    """The parallel sweep's Metropolis scan tests `pflip_window` proposals at
    once against the running state and applies the first acceptance; with a
    window of 1 it is the sequential walk itself.  Level 5 (batches of 30
    proposals) without repulsion accepts ~75% of the flips, so most batches
    hold many acceptances: windows 1, 7 and 256 give the same accepted counts
    and faces."""
    v, f, p = flip_case(5, 25, False)
    runs = {}
    for w in (1, 7, 256):
        gm = core.TriMesh(v, f)
        ge = core.EnergyManager(gm, to_core_params(core, p))
        t = ge.tuning
        t["pflip_window"] = w
        ge.tuning = t
        assert ge.tuning["pflip_window"] == w
        accs = [core.pflip(gm, ge, 0.2, seed=4321 + k) for k in range(3)]
        runs[w] = (accs, gm.fv_indices().copy())
        print(w, accs, flush=True)
    assert min(runs[1][0]) > 0.2 * 0.2 * len(f) * 3 // 2
    for w in (7, 256):
        assert runs[w][0] == runs[1][0]
        assert np.array_equal(runs[w][1], runs[1][1])


def test_pflip_topology_consistent_with_oracle(core):
    """After parallel sweeps the device topology, ring table and host mirror
    agree: properties, energy and gradient match an oracle mesh rebuilt from
    fv_indices (area/volume/curvature penalties off, so the reference values
    of the two managers do not enter)."""
    v, f, p = flip_case(3, 22, True)
    p.kappa_a = p.kappa_v = p.kappa_c = 0.0
    gm = core.TriMesh(v, f)
    ge = core.EnergyManager(gm, to_core_params(core, p))
    for k in range(4):
        core.pflip(gm, ge, 0.5, seed=77 + k)
    fv = gm.fv_indices()
    assert not np.array_equal(fv, f)
    om = O.Mesh(v, fv)
    oe = O.EnergyManager(om, p)
    assert rel(props_vec(ge.properties(gm)), oe.properties(om)) < 1e-10
    assert abs(ge.energy(gm) - oe.energy(om)) <= 1e-10 * abs(oe.energy(om))
    assert rel(ge.gradient(gm), oe.gradient(om)) < 1e-10


@pytest.mark.parametrize("kind", ["serial", "parallel"])
def test_exclusions_follow_flips(core, kind):
    """The cell-list query skips excluded candidates by per-vertex masks built
    for one topology; after every flip sweep the energy and gradient at
    exclusion level 2 match an oracle mesh rebuilt from the new faces."""
    v, f, p = flip_case(3, 31, True)
    p.rep_exclusion = 2
    p.kappa_a = p.kappa_v = p.kappa_c = 0.0
    gm = core.TriMesh(v, f)
    ge = core.EnergyManager(gm, to_core_params(core, p))
    fn = core.flip if kind == "serial" else core.pflip
    ge.energy(gm)
    for k in range(3):
        assert fn(gm, ge, 0.3, seed=5 + k) > 0
        om = O.Mesh(v, gm.fv_indices())
        oe = O.EnergyManager(om, p)
        e = oe.energy(om)
        assert abs(ge.energy(gm) - e) <= 1e-10 * abs(e), k
        assert rel(ge.gradient(gm), oe.gradient(om)) < 1e-10, k


def test_flip_ratio_errors(core):
    """flips.cpp:24-25 only constructs the error for flip; pflip raises
    (flips.cpp:84-85, std::range_error -> ValueError)."""
    v, f, p = flip_case(2, 23, False)
    gm = core.TriMesh(v, f)
    ge = core.EnergyManager(gm, to_core_params(core, p))
    with pytest.raises(ValueError):
        core.pflip(gm, ge, 1.5, seed=1)
    core.flip(gm, ge, 1.2, seed=1)
    check_manifold(gm.fv_indices(), gm.n_vertices(), gm.n_edges())


def test_pflip_statistics_vs_serial(core):
    """Both sweeps sample the same flip Markov chain; their mean acceptance
    per attempted flip and mean energy over 40 sweeps agree statistically."""
    v, f, p = flip_case(3, 24, False)
    stats = {}
    for kind in ("serial", "parallel"):
        gm = core.TriMesh(v, f)
        ge = core.EnergyManager(gm, to_core_params(core, p))
        accs, ens = [], []
        for k in range(40):
            fn = core.flip if kind == "serial" else core.pflip
            accs.append(fn(gm, ge, 0.2, seed=1000 + k))
            ens.append(ge.energy(gm))
        stats[kind] = (np.mean(accs[10:]), np.mean(ens[10:]))
    (a_s, e_s), (a_p, e_p) = stats["serial"], stats["parallel"]
    assert abs(a_p - a_s) < 0.35 * a_s, stats
    assert abs(e_p - e_s) < 0.05 * abs(e_s), stats


# ------------------------------------------------ finite differences -------
@pytest.mark.parametrize("bond", [0, 1])
def test_fd_gradient(core, bond):
    """core.gradient(mesh, estore, gradient, epsilon), the forward-difference
    gradient of bindings.cpp:25-50, on the reference's test_gradient.py
    setup (icosphere(3), kappas 1, fractions 0.8, lc0 1.15 l, lc1 0.85 l,
    a0 = mean face area, epsilon 1e-8): within 1e-4 of the analytic gradient
    (the reference test's bound) and, against the oracle's forward
    differences, within 4x their rounding noise: each difference quotient
    carries about 2 eps |E| / h of rounding, the two codes round their
    energies independently (a factor sqrt 2), and the norm over n components
    grows as sqrt n."""
    v, f = M.icosphere(3)
    om = O.Mesh(v, f)
    a, l = om.avg_tri_props()
    p = O.Params(kappa_b=1.0, kappa_a=1.0, kappa_v=1.0, kappa_c=1.0, kappa_t=1.0, area_frac=0.8,
                 volume_frac=0.8, curvature_frac=0.8, bond_type=bond, bond_lc0=1.15 * l, bond_lc1=0.85 * l,
                 bond_a0=a)
    oe = O.EnergyManager(om, p)
    gm = core.TriMesh(v, f)
    ge = core.EnergyManager(gm, to_core_params(core, p))
    fd = np.empty((gm.n_vertices(), 3))
    core.gradient(gm, ge, fd, 1.0e-8)
    assert np.abs(gm.points() - v).max() <= 1e-15    # each probe is undone by subtracting epsilon
    g_or = oe.fd_gradient(om, 1.0e-8)
    e_an, e_fd = rel(fd, ge.gradient(gm)), rel(fd, g_or)
    noise = 2.0 * np.finfo(float).eps * abs(oe.energy(om)) / 1.0e-8 * np.sqrt(2.0 * fd.size) / np.linalg.norm(g_or)
    print(f"bond {bond}: fd vs analytic {e_an:.2e}, fd vs oracle fd {e_fd:.2e} (noise {noise:.2e})", flush=True)
    assert e_an < 1e-4
    assert e_fd < 4.0 * noise


def bipyramid(n=32, h=0.3):
    """Closed bipyramid over an n-gon: the two apexes have valence n, the
    equator vertices 4.  Flipping an equator edge joins the two apexes."""
    t = 2 * np.pi * np.arange(n) / n
    v = np.concatenate([np.stack([np.cos(t), np.sin(t), np.zeros(n)], 1), [[0, 0, h], [0, 0, -h]]])
    i = np.arange(n)
    f = np.concatenate([np.stack([np.full(n, n), i, (i + 1) % n], 1),
                        np.stack([np.full(n, n + 1), (i + 1) % n, i], 1)]).astype(np.int32)
    return v, f


def max_valence(f, nv):
    e = np.unique(np.sort(f[:, [0, 1, 1, 2, 2, 0]].reshape((-1, 2)), axis=1), axis=0)
    return int(np.bincount(e.reshape(-1), minlength=nv).max())


def test_flips_keep_valence_within_ring_table(core):
    """A flip that would give a vertex more neighbours than the ring table
    holds (32) is refused, a deviation from the reference, which has no cap:
    on the 32-gon bipyramid the oracle accepts the flip of an equator edge
    with a zero uniform (both apexes reach valence 33), the HIP sweeps
    (serial one-workgroup and windowed, parallel) refuse it and every later
    evaluation still covers every vertex."""
    v, f = bipyramid()
    lbar = M.mean_edge_length(v, f)
    # area/volume/curvature penalties off: the oracle rebuilt from the flipped
    # faces takes its reference values from them
    p = base_params(lbar, kappa_a=0.0, kappa_v=0.0, kappa_c=0.0, kappa_t=1.0, kappa_r=0.0)
    for w in (0, 256):
        gm = core.TriMesh(v, f)
        ge = core.EnergyManager(gm, to_core_params(core, p))
        t = ge.tuning
        t["serial_window"] = w
        ge.tuning = t
        ne = gm.n_edges()
        prop = np.arange(ne, dtype=np.int32)
        acc, used = core.flip_injected(gm, ge, prop, np.zeros(ne))
        fv = gm.fv_indices()
        print("window", w, "accepted", acc, "used", used, "max valence", max_valence(fv, len(v)), flush=True)
        assert max_valence(fv, len(v)) <= 32
        om2 = O.Mesh(v, fv)
        oe2 = O.EnergyManager(om2, p)
        assert abs(ge.energy(gm) - oe2.energy(om2)) <= 1e-10 * abs(oe2.energy(om2))
    # the oracle (no cap) accepts flipping an equator edge of the original mesh
    om3 = O.Mesh(v, f)
    oe3 = O.EnergyManager(om3, p)
    ne = om3.n_edges()
    acc_o, _ = oe3.flip_serial_injected(om3, np.arange(ne, dtype=np.int32), np.zeros(ne))
    assert acc_o > 0 and max_valence(om3.fv_indices(), len(v)) > 32
    gm = core.TriMesh(v, f)
    ge = core.EnergyManager(gm, to_core_params(core, p))
    for k in range(4):
        core.pflip(gm, ge, 1.0, seed=50 + k)
        assert max_valence(gm.fv_indices(), len(v)) <= 32
