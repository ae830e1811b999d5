"""GPU: replica groups (trimem_amd.core.Replicas, the tm_replicas_* ABI).

Pinned by per-replica equality with single-mesh runs:
  * one batched evaluation of meshes of different sizes, parameters,
    neighbour lists and a boundary mesh gives, for every replica, the same
    energy and gradient bits as EnergyManager.energy / gradient on that mesh
    alone (FP64 and mixed);
  * grouped HMC steps with device draws give the same positions and
    acceptance counts as DeviceHMC on each mesh alone with the same seed and
    step numbers;
  * replaying each step from a captured hipGraph gives the same bits as
    direct launches;
  * replica exchange: every swap decision equals the one recomputed on the
    host from the replicas' bending B, the ladder and the uniform of the
    counter-based stream (restated below from its definition), and after a
    swap each replica's energy changes by (new kappa_b - old kappa_b) * B.
"""
import ctypes

import numpy as np
import pytest

from trimem_amd import meshes as M

pytestmark = pytest.mark.gpu

MASK = (1 << 64) - 1


def splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def rng_unit(seed, stream, i):
    b = splitmix64((splitmix64(seed ^ ((stream * 0xD1B54A32D192ED03) & MASK)) + i) & MASK)
    return (b >> 11) * (1.0 / 9007199254740992.0)


@pytest.fixture(scope="module")
def core():
    import trimem_amd.core as m
    assert m.device_count() > 0, "no GPU visible"
    return m


def params(m, lbar, kappa_b=1.0, vfrac=0.9, n_search="cell-list", excl=0, kappa_r=1.0):
    e = m.EnergyParams()
    e.kappa_b, e.kappa_a, e.kappa_v, e.kappa_c, e.kappa_t, e.kappa_r = kappa_b, 1e3, 1e3, 1.0, 10.0, kappa_r
    e.area_frac, e.volume_frac, e.curvature_frac = 1.0, vfrac, 1.0
    b = m.BondParams()
    b.type, b.r, b.lc0, b.lc1 = m.BondType.Edge, 2, 1.25 * lbar, 0.75 * lbar
    e.bond_params = b
    r = m.SurfaceRepulsionParams()
    r.n_search, r.rlist, r.lc1, r.r, r.exclusion_level = n_search, 1.2 * lbar, 1.2 * lbar, 2, excl
    e.repulse_params = r
    return e


def membrane(core, level, seed, precision, **kw):
    v, f, lbar = M.jittered_icosphere(level, seed)
    mesh = core.TriMesh(v, f)
    est = core.EnergyManager(mesh, params(core, lbar, **kw))
    est.precision = precision
    return mesh, est


def mixed_set(core, precision):
    out = [membrane(core, 2, 1, precision),
           membrane(core, 4, 2, precision, kappa_b=30.0, vfrac=0.8),
           membrane(core, 3, 3, precision, n_search="verlet-list", excl=1),
           membrane(core, 3, 4, precision, excl=2)]
    v, f = M.tube(length=1.0, radius=0.5, segments=24)
    lbar = M.mean_edge_length(v, f)
    tube = core.TriMesh(v, f)
    est = core.EnergyManager(tube, params(core, lbar))
    est.precision = precision
    out.append((tube, est))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_eval_equals_single(core, precision):
    reps = mixed_set(core, precision)
    R = core.Replicas([m for m, _ in reps], [e for _, e in reps])
    R.eval()
    g = R.globals()
    grads = [R.gradient(i) for i in range(len(reps))]
    for i, (m, e) in enumerate(reps):
        e_single = e.energy(m)
        g_single = e.gradient(m)
        print(precision, i, m.n_vertices(), g[i, 6], e_single, flush=True)
        assert g[i, 6] == e_single
        assert np.array_equal(grads[i], g_single)


def test_grouped_hmc_equals_single(core):
    seeds = [11, 22, 33]
    steps = 5
    opts = dict(mass=1.0, time_step=2e-3, temperature=1.0, num_integration_steps=8)
    reps = [membrane(core, 2, 5, "mixed"), membrane(core, 3, 6, "mixed"), membrane(core, 3, 7, "mixed", excl=2)]
    R = core.Replicas([m for m, _ in reps], [e for _, e in reps])
    R.hmc_setup(seeds=seeds, **opts)
    R.hmc_run(steps)
    st = R.stats()
    xs = [m.points().copy() for m, _ in reps]

    single = [membrane(core, 2, 5, "mixed"), membrane(core, 3, 6, "mixed"), membrane(core, 3, 7, "mixed", excl=2)]
    for i, (m, e) in enumerate(single):
        d = core.DeviceHMC(m, e)
        for k in range(steps):
            d.step(mass=opts["mass"], time_step=opts["time_step"], temperature=opts["temperature"],
                   num_integration_steps=opts["num_integration_steps"], seed=seeds[i], step=k)
        r = d.result()
        print(i, "accepted grouped", st["accepted"][i], "single", r["n_accepted"], flush=True)
        assert st["accepted"][i] == r["n_accepted"] and st["steps"][i] == steps
        assert np.array_equal(m.points(), xs[i])


def test_graph_replay_equals_direct(core):
    out = []
    for use_graph in (False, True):
        reps = [membrane(core, 3, 8, "fp64"), membrane(core, 2, 9, "fp64")]
        R = core.Replicas([m for m, _ in reps], [e for _, e in reps])
        R.hmc_setup(mass=1.0, time_step=2e-3, temperature=1.0, seeds=[5, 6], num_integration_steps=6)
        R.set_ladder([1.0, 2.0], seed=3)
        R.hmc_run(6, swap_every=2, use_graph=use_graph)
        st = R.stats()
        out.append(([m.points().copy() for m, _ in reps], st["accepted"].copy(), st["slot_of"].copy()))
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def test_exchange_rule(core):
    # close rungs (swaps mostly accepted) and one wide rung (mostly rejected)
    ladder = [10.0, 10.05, 10.1, 11.0]
    seed = 1234
    reps = [membrane(core, 3, 20 + i, "fp64", kappa_b=ladder[i]) for i in range(4)]
    R = core.Replicas([m for m, _ in reps], [e for _, e in reps])
    R.hmc_setup(mass=1.0, time_step=1e-3, temperature=1.0, seeds=[1, 2, 3, 4], num_integration_steps=4)
    R.set_ladder(ladder, seed=seed)
    R.hmc_run(1)
    decided = 0
    for rd in range(6):
        st = R.stats()
        slot_of = st["slot_of"].copy()
        owner = np.argsort(slot_of)
        B = st["props"][:, 3].copy()
        R.eval(gradient=False)
        e_before = R.globals()[:, 6].copy()
        R.swap()
        st2 = R.stats()
        expect = owner.copy()
        for s in range(rd % 2, 3, 2):
            a, b = owner[s], owner[s + 1]
            dE = (ladder[s + 1] - ladder[s]) * (B[a] - B[b])
            alpha = 1.0 if dE <= 0 else np.exp(-dE)
            if rng_unit(seed, 5, rd * 4 + s) <= alpha:
                expect[s], expect[s + 1] = b, a
            decided += 1
        print(rd, "B", B, "owner", owner, "->", np.argsort(st2["slot_of"]), flush=True)
        assert np.array_equal(np.argsort(st2["slot_of"]), expect)
        R.eval(gradient=False)
        e_after = R.globals()[:, 6]
        kb_before = np.array(ladder)[slot_of]
        kb_after = np.array(ladder)[st2["slot_of"]]
        assert np.allclose(e_after - e_before, (kb_after - kb_before) * B, rtol=1e-12, atol=1e-9)
        R.hmc_run(1)
    assert decided == 9
    assert st2["swaps_tried"].sum() == 9
    assert 0 < st2["swaps_accepted"].sum() < 9    # both outcomes exercised


SHAPES = [dict(rep_threads=64, rep_batch=1, pass2_threads=64, rep_fused=0, pair_cache=1, pflip_window=256,
               serial_window=256),
          dict(rep_threads=512, rep_batch=8, pass2_threads=512, rep_fused=1, pair_cache=0, pflip_window=1,
               serial_window=0),
          dict(rep_threads=128, rep_batch=2, pass2_threads=128, rep_fused=0, pair_cache=0, pflip_window=64,
               serial_window=7),
          dict(rep_threads=256, rep_batch=1, pass2_threads=256, rep_fused=1, pair_cache=1, pflip_window=256,
               serial_window=128),
          dict(rep_threads=256, rep_batch=4, pass2_threads=64, rep_fused=0, pair_cache=1, pflip_window=256,
               serial_window=256)]


def vertex_properties(mesh, est):
    """tm_energy_vertex_properties: the (N, 6) per-vertex properties."""
    from trimem_amd import capi
    fn = capi.lib().tm_energy_vertex_properties
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    out = np.zeros((mesh.n_vertices(), 6))
    capi.check(fn(est._handle(), mesh._handle(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_launch_shapes_change_no_bits(core, precision):
    """Every repulsion / pass-2 launch shape, the repulsion query inside or
    apart from the properties launch, and the autotuner's choice give the
    default shape's energy and gradient bits (exclusion 0 so that every vertex
    has repulsion pairs)."""
    mesh, est = membrane(core, 4, 31, precision, excl=0)
    e0, g0, p0 = est.energy(mesh), est.gradient(mesh), vertex_properties(mesh, est)
    assert np.all(p0[:, 5] > 0)
    for sh in SHAPES:
        est.tuning = sh
        assert est.tuning == sh
        assert est.energy(mesh) == e0 and np.array_equal(est.gradient(mesh), g0)
        assert np.array_equal(vertex_properties(mesh, est), p0)
    best = est.autotune(mesh)
    print(precision, "autotuned", best, flush=True)
    assert best["rep_threads"] in (64, 128, 256, 512) and best["rep_batch"] in (1, 2, 4, 8)
    assert est.energy(mesh) == e0 and np.array_equal(est.gradient(mesh), g0)


def test_grouped_shapes_and_autotune(core):
    reps = mixed_set(core, "mixed")
    R = core.Replicas([m for m, _ in reps], [e for _, e in reps])
    R.eval()
    ref = R.globals().copy(), [R.gradient(i) for i in range(len(reps))]
    for sh in SHAPES:
        R.tuning = sh
        R.eval()
        assert np.array_equal(R.globals(), ref[0])
        assert all(np.array_equal(R.gradient(i), ref[1][i]) for i in range(len(reps)))
    best = R.autotune()
    print("group autotuned", best, flush=True)
    R.eval()
    assert np.array_equal(R.globals(), ref[0])


def test_grouped_pair_cache_is_bit_identical(core):
    """Replica groups replay recorded candidate lists per replica (device
    recording flags, per-replica displacement checks); a group with the
    cache and one without agree bit for bit through HMC trajectories, a
    flip on one replica's mesh (its exclusions change) and a list rebuild.
    The Verlet-list replica takes the full query in both."""
    def group(cache):
        reps = mixed_set(core, "mixed")
        R = core.Replicas([m for m, _ in reps], [e for _, e in reps])
        t = R.tuning
        t["pair_cache"] = cache
        R.tuning = t
        R.hmc_setup(mass=1.0, time_step=2e-3, temperature=1.0, seeds=list(range(1, len(reps) + 1)),
                    num_integration_steps=6)
        return R, reps

    on, reps_on = group(1)
    off, reps_off = group(0)
    assert on.tuning["pair_cache"] == 1 and off.tuning["pair_cache"] == 0

    def same(what):
        on.eval()
        off.eval()
        assert np.array_equal(on.globals(), off.globals()), what
        for i in range(len(reps_on)):
            assert np.array_equal(on.gradient(i), off.gradient(i)), (what, i)
            assert np.array_equal(reps_on[i][0].points(), reps_off[i][0].points()), (what, i)

    same("start")
    on.hmc_run(4)
    off.hmc_run(4)
    same("after hmc")
    for (m, e) in (reps_on[1], reps_off[1]):
        assert core.flip(m, e, 0.3, seed=17) > 0
    same("after flip")
    on.hmc_run(3)
    off.hmc_run(3)
    same("after flip and hmc")
    on.update_repulsion()
    off.update_repulsion()
    on.hmc_run(3)
    off.hmc_run(3)
    same("after rebuild and hmc")


def test_member_rebuild_refreshes_group(core):
    """A member's own update_repulsion (the reference callback's path,
    evaluators.py:85) swaps in a new neighbour list; the group notices and
    evaluates with the member's current list: its energies and gradients stay
    bit-equal to the single-mesh calls, also after the member's list buffers
    grow, and a captured HMC step is recaptured against the new table."""
    reps = [membrane(core, 3, 41, "fp64"), membrane(core, 3, 42, "fp64", excl=2)]
    R = core.Replicas([m for m, _ in reps], [e for _, e in reps])
    R.hmc_setup(seeds=[1, 2], mass=1.0, time_step=1e-3, temperature=1.0, num_integration_steps=4)
    R.hmc_run(2, use_graph=True)
    m0, e0 = reps[0]
    x = m0.points()
    x *= 1.5                                     # a larger box: the cell grid grows
    e0.update_repulsion(m0)
    R.eval()
    g = R.globals()
    for i, (m, e) in enumerate(reps):
        assert g[i, 6] == e.energy(m), i
        assert np.array_equal(R.gradient(i), e.gradient(m)), i
    R.hmc_run(2, use_graph=True)                 # replays against the refreshed table
    for i, (m, e) in enumerate(reps):
        R.eval()
        assert R.globals()[i, 6] == e.energy(m), i


def test_grouped_pflip_equals_single(core):
    """Replicas.pflip: every replica's accepted flips, faces and the energy
    and gradient after them equal pflip on that mesh alone with the same
    seed (meshes of different sizes, so different batch sizes and rounds,
    both list types, an open tube), over two sweeps; and a grouped
    evaluation after the sweeps sees the new topologies."""
    seeds = [101, 202, 303, 404, 505]
    a = mixed_set(core, "mixed")
    b = mixed_set(core, "mixed")
    R = core.Replicas([m for m, _ in a], [e for _, e in a])
    R.eval()
    for sweep in range(2):
        got = R.pflip(0.2, [s + sweep for s in seeds])
        want = [core.pflip(m, e, 0.2, seed=s + sweep) for (m, e), s in zip(b, seeds)]
        assert got == want
        assert sum(got) > 0
        for (ma, ea), (mb, eb) in zip(a, b):
            assert np.array_equal(ma.fv_indices(), mb.fv_indices())
            assert ea.energy(ma) == eb.energy(mb)
    R.eval()
    g = R.globals()
    for i, ((ma, ea), (mb, eb)) in enumerate(zip(a, b)):
        assert g[i, 6] == eb.energy(mb)
    with pytest.raises(Exception):
        R.pflip(1.5, seeds)


def test_grouped_pflip_uses_the_ladder(core):
    """With a kappa_b ladder, each replica flips with the kappa_b of the slot
    it holds: the same accepted flips and faces as pflip on a mesh whose
    own kappa_b is that slot's value (the group's meshes were created with
    kappa_b 1, so ignoring the ladder would decide differently)."""
    ladder = [5.0, 20.0, 60.0, 150.0]
    seeds = [71, 72, 73, 74]
    a = [membrane(core, 3, 40 + i, "fp64", kappa_b=1.0) for i in range(4)]
    b = [membrane(core, 3, 40 + i, "fp64", kappa_b=ladder[i]) for i in range(4)]
    R = core.Replicas([m for m, _ in a], [e for _, e in a])
    R.hmc_setup(mass=1.0, time_step=1e-4, temperature=1.0, seeds=[1, 2, 3, 4], num_integration_steps=2)
    R.set_ladder(ladder, seed=5)
    slot_of = R.stats()["slot_of"]
    assert list(slot_of) == [0, 1, 2, 3]
    got = R.pflip(0.5, seeds)
    want = [core.pflip(m, e, 0.5, seed=s) for (m, e), s in zip(b, seeds)]
    assert got == want
    for (ma, _), (mb, _) in zip(a, b):
        assert np.array_equal(ma.fv_indices(), mb.fv_indices())
