"""GPU: the pair cache of the cell-list repulsion query.

The cache replays, per vertex and lane, the candidates a recording launch met
within lc1 + skin, while every vertex has moved less than skin / 2 since the
recording and the vertex is in the cell it was recorded in; otherwise the
query runs in full (and records again).  Each lane then adds the same terms
in the same order as the full query, so an energy manager with the cache and
one without it must agree bit for bit on energy, gradient and per-vertex
properties through every kind of change: small moves (replay, with some
vertices crossing cell boundaries), large moves (recording forced by the
displacement), flips (exclusion may change), neighbour-list rebuilds, and
device-resident HMC trajectories.
"""
import ctypes

import numpy as np
import pytest

from trimem_amd import meshes as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    import trimem_amd.core as m
    assert m.device_count() > 0, "no GPU visible"
    return m


def params(m, lbar, excl, rc=1.5):
    e = m.EnergyParams()
    e.kappa_b, e.kappa_a, e.kappa_v, e.kappa_c, e.kappa_t, e.kappa_r = 30.0, 1e3, 1e3, 1.0, 10.0, 1e2
    e.area_frac, e.volume_frac, e.curvature_frac = 1.0, 0.9, 1.0
    b = m.BondParams()
    b.type, b.r, b.lc0, b.lc1 = m.BondType.Edge, 2, 1.25 * lbar, 0.75 * lbar
    e.bond_params = b
    r = m.SurfaceRepulsionParams()
    r.n_search, r.rlist, r.lc1, r.r, r.exclusion_level = "cell-list", rc * lbar, rc * lbar, 2, excl
    e.repulse_params = r
    return e


def vertex_properties(mesh, est):
    from trimem_amd import capi
    fn = capi.lib().tm_energy_vertex_properties
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    out = np.zeros((mesh.n_vertices(), 6))
    capi.check(fn(est._handle(), mesh._handle(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def pair(core, mesh, lbar, excl, precision, rc=1.5):
    on = core.EnergyManager(mesh, params(core, lbar, excl, rc))
    off = core.EnergyManager(mesh, params(core, lbar, excl, rc))
    for e in (on, off):
        e.precision = precision
    t = off.tuning
    t["pair_cache"] = 0
    off.tuning = t
    assert on.tuning["pair_cache"] == 1 and off.tuning["pair_cache"] == 0
    return on, off


def same(mesh, on, off, what):
    e1, e0 = on.energy(mesh), off.energy(mesh)
    g1, g0 = on.gradient(mesh), off.gradient(mesh)
    p1, p0 = vertex_properties(mesh, on), vertex_properties(mesh, off)
    print(what, e0, float(p0[:, 5].sum()), flush=True)
    assert e1 == e0, what
    assert np.array_equal(g1, g0), what
    assert np.array_equal(p1, p0), what
    return p0


@pytest.mark.parametrize("excl", [0, 1, 2])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cache_is_bit_identical(core, excl, precision):
    v, f, lbar = M.jittered_icosphere(4, seed=40 + excl)
    mesh = core.TriMesh(v, f)
    on, off = pair(core, mesh, lbar, excl, precision)
    rng = np.random.default_rng(excl)
    rep = [same(mesh, on, off, "start")[:, 5].sum()]
    for k in range(16):
        x = mesh.points()
        kind = k % 4
        if kind in (0, 1):
            # small moves: replay (max |dx| ~ 0.05 lbar < 0.25 * 0.2 * 1.5 lbar);
            # a few percent of the vertices cross a cell boundary
            x[:] = x + rng.normal(0, 0.01 * lbar, x.shape)
        elif kind == 2:
            # large move: recording forced by the displacement
            x[:] = x + rng.normal(0, 0.2 * lbar, x.shape)
        else:
            if k % 8 == 3:
                assert core.flip(mesh, on, 0.2, seed=k) > 0
            else:
                on.update_repulsion(mesh)
                off.update_repulsion(mesh)
        rep.append(same(mesh, on, off, f"step {k} kind {kind}")[:, 5].sum())
    assert max(rep) > 0    # the sequence reaches states with repulsion pairs


def test_cache_with_boundary_vertices(core):
    """Open tube: boundary vertices take the generic walker; they still enter
    the displacement bound as candidates of ring vertices."""
    v, f = M.tube(length=1.0, radius=0.5, segments=24)
    lbar = M.mean_edge_length(v, f)
    mesh = core.TriMesh(v, f)
    on, off = pair(core, mesh, lbar, 0, "mixed", rc=1.2)
    rng = np.random.default_rng(3)
    same(mesh, on, off, "start")
    for k in range(6):
        x = mesh.points()
        x[:] = x + rng.normal(0, (0.01 if k % 2 else 0.2) * lbar, x.shape)
        same(mesh, on, off, f"tube {k}")


def test_cache_in_device_hmc(core):
    """Device HMC trajectories (11 force evaluations each, the list rebuilt
    every other step) with and without the cache end bit-identical."""
    v, f, lbar = M.jittered_icosphere(4, seed=5)
    runs = []
    for cache in (1, 0):
        mesh = core.TriMesh(v, f)
        est = core.EnergyManager(mesh, params(core, lbar, 2))
        est.precision = "mixed"
        t = est.tuning
        t["pair_cache"] = cache
        est.tuning = t
        d = core.DeviceHMC(mesh, est)
        out = []
        for s in range(8):
            if s % 2 == 0:
                est.update_repulsion(mesh)
            d.step(mass=1.0, time_step=2e-4, num_integration_steps=10, temperature=1e-3, seed=9, step=s)
            r = d.result()
            out.append((r["accepted"], r["dh"], r["energy1"]))
        runs.append((mesh.points().copy(), out))
        print("cache", cache, [o[0] for o in out], flush=True)
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][1] == runs[1][1]
    assert any(o[0] for o in runs[0][1])
