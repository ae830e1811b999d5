"""GPU: the parallel flip sweep's decisions used the right energy differences.

Each round, proposals claim their flip patch (the edges of every outgoing
halfedge of the patch's four vertices and the next edges of those
halfedges); only a proposal that kept all of its patch edges is evaluated
and may flip.  The Metropolis scan then adds each accepted flip's patch
difference (properties after minus before, evaluated on the topology at the
start of the round) to its running global properties.  Those differences are
right only if no two flips accepted in one round share a patch edge, so the
running properties at the end of a sweep (tm_energy_pflip_state) must equal
the properties of the final topology, evaluated afresh.  Large batches (E/8
and E/16 proposals per round) make many proposals conflict, so a claim that
missed an edge would show here.

Area, volume, curvature, bending and tethering change only at the patch's
four vertices, so they must match at every exclusion level.  The repulsion
property must match only without exclusions: the patch difference is the
reference's edge_vertex_properties (flip_utils.cpp:19-41), the sum over the
four patch vertices, so when a flip changes whether a pair is excluded it
counts the change on the patch vertex's side only (and at exclusion level 2
not at all for pairs of two vertices outside the patch).  The reference's
running Metropolis energy drifts the same way; at exclusion level 2 the test
checks the other five properties.
"""
import ctypes

import numpy as np
import pytest

from trimem_amd import capi
from trimem_amd import meshes as M

from test_gpu_pair_cache import params
from test_gpu_parity import props_vec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    import trimem_amd.core as m
    assert m.device_count() > 0, "no GPU visible"
    return m


def sweep_state(e):
    st = (ctypes.c_double * 8)()
    assert capi.lib().tm_energy_pflip_state(ctypes.c_void_p(e._handle()), st) == 0
    return np.array(st[:])


def assert_matches(core, st, m, e, acc, excl):
    assert int(st[7]) == acc
    pr = props_vec(e.properties(m))
    for k in range(6 if excl == 0 else 5):
        assert abs(st[k] - pr[k]) <= 1e-10 * abs(pr[k]) + 1e-12, (k, st[k], pr[k])
    if excl == 0:
        en = e.energy(m)
        assert abs(st[6] - en) <= 1e-10 * abs(en), (st[6], en)


@pytest.mark.parametrize("excl,div", [(0, 8), (2, 16)])
def test_pflip_running_state_is_the_final_topology(core, excl, div):
    v, f, lbar = M.jittered_icosphere(5, seed=17, rel_sigma=0.02)
    m = core.TriMesh(v, f)
    e = core.EnergyManager(m, params(core, lbar, excl))
    e.precision = "fp64"
    f0 = m.fv_indices().copy()
    total = 0
    for k in range(3):
        acc = core.pflip(m, e, 0.5, seed=100 + k, batch=m.n_edges() // div)
        assert acc > 0
        total += acc
        assert_matches(core, sweep_state(e), m, e, acc, excl)
    assert not np.array_equal(m.fv_indices(), f0)
    print(excl, div, total, flush=True)


def test_grouped_pflip_running_state_is_the_final_topology(core):
    ms, es = [], []
    for level, seed in ((4, 3), (5, 4), (3, 5)):
        v, f, lbar = M.jittered_icosphere(level, seed=seed, rel_sigma=0.02)
        m = core.TriMesh(v, f)
        e = core.EnergyManager(m, params(core, lbar, 2))
        e.precision = "fp64"
        ms.append(m)
        es.append(e)
    R = core.Replicas(ms, es)
    for k in range(2):
        accs = R.pflip(0.4, [50 + 10 * k + i for i in range(len(ms))])
        for m, e, acc in zip(ms, es, accs):
            assert_matches(core, sweep_state(e), m, e, acc, 2)
