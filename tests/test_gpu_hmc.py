"""GPU: the device-resident HMC step (trimem_amd.core.DeviceHMC and
trimem_amd.mc.hmc.DeviceMeshHMC).

Pinned by:
  * the reference sampler test's trajectory ``ref_hmc``
    (tests/test_sampling.py:32-62): host-RNG mode on the icosahedron with
    E = x.x / 2, end point within 1e-14;
  * equality with the drop-in path: MeshHMC over EnergyEvaluators over
    trimem_amd.core (the reference's own wiring, util.py:174-229; its state
    decoupled from the mesh view as after a first acceptance) and
    DeviceMeshHMC from the same numpy seed give bit-identical positions and
    the same acceptances, FP64 and mixed precision;
  * the Metropolis rule re-checked on the host from each step's returned
    dh and uniform; rejected steps restore the start positions bit for bit;
  * device draws: E[K] = 3N T / 2 for the momenta (chi-square mean),
    x ~ N(0, T) for the harmonic chain, the Verlet energy error falling as
    dt^2 on a membrane, determinism per (seed, step).
"""
import json
import os
from collections import Counter

import numpy as np
import pytest

from trimem_amd import meshes as M

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


@pytest.fixture(scope="module")
def core():
    import trimem_amd.core as m
    assert m.device_count() > 0, "no GPU visible"
    return m


@pytest.fixture(scope="module")
def mc():
    from trimem_amd.mc import hmc, mesh, evaluators
    return hmc, mesh, evaluators


def membrane_params(m, lbar, n_search="cell-list"):
    e = m.EnergyParams()
    e.kappa_b, e.kappa_a, e.kappa_v, e.kappa_c, e.kappa_t, e.kappa_r = 1.0, 1e3, 1e3, 1.0, 10.0, 1.0
    e.area_frac, e.volume_frac, e.curvature_frac = 1.0, 0.9, 1.0
    b = m.BondParams()
    b.type, b.r, b.lc0, b.lc1 = m.BondType.Edge, 2, 1.25 * lbar, 0.75 * lbar
    e.bond_params = b
    r = m.SurfaceRepulsionParams()
    r.n_search, r.rlist, r.lc1, r.r, r.exclusion_level = n_search, 1.2 * lbar, 1.2 * lbar, 2, 0
    e.repulse_params = r
    return e


def membrane(core, precision, level=3, seed=11):
    v, f, lbar = M.jittered_icosphere(level, seed)
    from trimem_amd.mc.mesh import Mesh
    mesh = Mesh(v, f)
    est = core.EnergyManager(mesh.trimesh, membrane_params(core, lbar))
    est.precision = precision
    return mesh, est


def test_device_hmc_reference_trajectory(core, mc):
    hmc, mesh_mod, _ = mc
    np.random.seed(42)
    p, c = M.icosahedron()
    mesh = mesh_mod.Mesh(p, c)
    d = hmc.DeviceMeshHMC(mesh, None, counter=hmc.get_step_counters(), model="harmonic",
                          options={"info_step": 10, "time_step": 1.0e-1})
    d.run(10)
    x = mesh.x
    print("ref_hmc distance", np.linalg.norm(x - np.array(GOLD["ref_hmc"]["x"])), flush=True)
    assert np.linalg.norm(x - np.array(GOLD["ref_hmc"]["x"])) < GOLD["ref_hmc"]["tolerance"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_hmc_equals_dropin_path(core, mc, precision):
    hmc, _, ev = mc
    opts = {"time_step": 1.0e-3, "num_integration_steps": 10, "info_step": 1000}
    steps = 10

    mesh_a, est_a = membrane(core, precision)
    funcs = ev.EnergyEvaluators(mesh_a, est_a, None,
                                {"info_step": 0, "output_step": 0, "refresh_step": 0})
    np.random.seed(2024)
    host = hmc.MeshHMC(mesh_a, funcs.fun, funcs.grad, counter=Counter(), options=opts)
    # MeshHMC starts from the mesh's points view, which the evaluators
    # overwrite during the first trajectories (hmc.py:221, evaluators.py:105);
    # a copy gives the chain the reference has after its first acceptance
    host.x = mesh_a.x.copy()
    host.run(steps)

    mesh_b, est_b = membrane(core, precision)
    np.random.seed(2024)
    dev = hmc.DeviceMeshHMC(mesh_b, est_b, counter=Counter(), options=opts, rng="host")
    dev.run(steps)

    xa, xb = mesh_a.x.copy(), mesh_b.x.copy()
    print(precision, "accepted host", host.acc, "device", dev.acc, "max |dx|", np.abs(xa - xb).max(), flush=True)
    assert host.acc == dev.acc
    assert 0 < host.acc < steps    # both branches taken
    assert np.array_equal(xa, xb)


def test_metropolis_rule_and_restore(core):
    mesh, est = membrane(core, "fp64")
    d = core.DeviceHMC(mesh.trimesh, est)
    n = mesh.trimesh.n_vertices()
    x_prev = mesh.x.copy()
    rng = np.random.default_rng(5)
    seen = set()
    for k, (dt, u) in enumerate([(1e-3, 0.5), (3e-2, 0.5), (1e-3, 0.0), (5e-2, 0.999), (2e-3, 0.3),
                                 (1e-1, 0.01)]):
        p = rng.normal(size=(n, 3))
        d.step(mass=1.0, time_step=dt, num_integration_steps=6, temperature=1.0, momenta=p, uniform=u)
        r = d.result()
        alpha = 1.0 if r["dh"] <= 0 else np.exp(-r["dh"])
        print(k, dt, u, r["dh"], r["accepted"], flush=True)
        assert r["uniform"] == u
        assert r["accepted"] == (u <= alpha)
        x = mesh.x.copy()
        if r["accepted"]:
            assert not np.array_equal(x, x_prev)
            # host-facing evaluation sees the device state
            assert est.energy(mesh.trimesh) == r["energy1"]
        else:
            assert np.array_equal(x, x_prev)
            assert est.energy(mesh.trimesh) == r["energy0"]
        seen.add(r["accepted"])
        x_prev = x
    assert seen == {True, False}
    assert d.result()["n_steps"] == 6


def test_device_momenta_chi_square(core, mc):
    """L = 0: no motion, always accepted; K0 = p.p / 2 with p ~ N(0, m T)
    has mean 3 N m T / (2 m) = 3 N T / 2."""
    hmc, mesh_mod, _ = mc
    p, c = M.icosahedron()
    mesh = mesh_mod.Mesh(p, c)
    d = core.DeviceHMC(mesh.trimesh, None, "harmonic")
    T, k = 1.7, []
    for s in range(1500):
        d.step(mass=2.0, time_step=0.1, num_integration_steps=0, temperature=T, seed=9, step=s)
        r = d.result()
        assert r["accepted"] and r["dh"] == 0.0
        k.append(r["kinetic0"])
    k = np.array(k)
    mean_ref = 3 * 12 * T / 2
    se = mean_ref * np.sqrt(2.0 / 36) / np.sqrt(len(k))
    print("K mean", k.mean(), "expected", mean_ref, "se", se, flush=True)
    assert abs(k.mean() - mean_ref) < 5 * se
    assert np.array_equal(mesh.x, p)


def test_device_harmonic_equipartition(core, mc):
    """HMC on E = x.x / 2 at temperature T samples x ~ N(0, T)."""
    hmc, mesh_mod, _ = mc
    p, c = M.icosahedron()
    mesh = mesh_mod.Mesh(p, c)
    T = 0.6
    d = hmc.DeviceMeshHMC(mesh, None, counter=Counter(), model="harmonic", rng="device", seed=123,
                          options={"time_step": 0.25, "num_integration_steps": 7, "initial_temperature": T,
                                   "info_step": 0})
    xs = []
    for _ in range(800):
        d.step()
        xs.append(mesh.x.copy())
    xs = np.array(xs[100:])
    var = xs.var()
    print("var", var, "mean", xs.mean(), "acc", d.acc / 800, flush=True)
    assert abs(xs.mean()) < 0.1
    assert abs(var / T - 1.0) < 0.1
    assert d.acc > 0.8 * 800


def test_device_rng_determinism(core, mc):
    hmc, _, _ = mc
    out = []
    for seed in (1, 1, 2):
        mesh, est = membrane(core, "mixed")
        d = hmc.DeviceMeshHMC(mesh, est, counter=Counter(), rng="device", seed=seed,
                              options={"time_step": 2e-3, "num_integration_steps": 5, "info_step": 0})
        d.run(3)
        out.append(mesh.x.copy())
    assert np.array_equal(out[0], out[1])
    assert not np.array_equal(out[0], out[2])


def test_verlet_energy_error_order(core):
    """Same momenta (device draw keyed by seed and step), trajectory time
    fixed at 0.04: halving dt quarters |dh| (T = 0.01 keeps the trajectory
    in the smooth regime of this stiff membrane)."""
    dh = {}
    for dt, L in ((4e-3, 10), (2e-3, 20), (1e-3, 40)):
        vals = []
        for s in range(3):
            mesh, est = membrane(core, "fp64", seed=21)
            d = core.DeviceHMC(mesh.trimesh, est)
            d.step(mass=1.0, time_step=dt, num_integration_steps=L, temperature=0.01, seed=77, step=s)
            vals.append(abs(d.result()["dh"]))
        dh[dt] = np.mean(vals)
    r1, r2 = dh[4e-3] / dh[2e-3], dh[2e-3] / dh[1e-3]
    print("dh", dh, "ratios", r1, r2, flush=True)
    assert 3.5 < r1 < 4.6 and 3.5 < r2 < 4.6


def test_combined_chain_with_device_hmc(core, mc):
    """The reference's combined chain (util.py:174-237) with the device HMC,
    parallel flips and the evaluators' callback (repulsion refresh,
    continuation) keeps a valid membrane."""
    hmc, _, ev = mc
    mesh, est = membrane(core, "mixed", seed=4)
    ne = mesh.trimesh.n_edges()
    np.random.seed(3)
    counter = hmc.get_step_counters()
    funcs = ev.EnergyEvaluators(mesh, est, None, {"info_step": 0, "output_step": 0, "refresh_step": 2})
    h = hmc.DeviceMeshHMC(mesh, est, counter=counter, rng="device", seed=5,
                          options={"time_step": 2e-3, "num_integration_steps": 10, "info_step": 0})
    fl = hmc.MeshFlips(mesh, est, counter=counter, options={"flip_type": "parallel", "flip_ratio": 0.1,
                                                            "info_step": 0})
    chain = hmc.MeshMonteCarlo(h, fl, counter=counter, callback=funcs.callback)
    chain.run(12)
    assert counter["move"] + counter["flip"] == 12
    assert mesh.trimesh.n_edges() == ne
    e = est.energy(mesh.trimesh)
    assert np.isfinite(e) and np.isfinite(mesh.x).all()
