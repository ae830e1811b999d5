"""The benchmark configurations as inputs: meshes, parameters and ladders.

Every configuration is deterministic (numpy default_rng from a fixed seed),
so the benchmark and the parity tests build identical inputs.  Parameter
dicts use the field names of tm_params / the oracle's Params; core_params()
turns one into trimem_amd.core.EnergyParams.

  single(level)     one icosphere of the given level (level 4: the
                    CPU-runnable plain configuration; level 7: the single
                    large mesh), jittered by N(0, (0.01 l)^2), kappa_b=30,
                    kappa_a=kappa_v=1e6, kappa_c=0, kappa_t=1e5, kappa_r=1e3,
                    area fraction 1, volume fraction 0.8, Edge tether
                    (r=2, lc0=1.25 l, lc1=0.75 l), cell-list repulsion with
                    exclusion level 2 and rc = rlist = 1.5 l, where l is the
                    mean edge length of the unjittered mesh
  replicas()        64 meshes of levels 3-6 drawn from a seed, volume
                    fractions U[0.6, 0.95], kappa_b log-uniform on [10, 100]
                    sorted into a tempering ladder
  combined()        32 level-7 meshes jittered from their own seeds, volume
                    fractions U[0.55, 0.9], kappa_b log-uniform on [20, 80]
"""
import numpy as np

from . import meshes

SINGLE_SEED = 20261203
REPLICA_SEED = 20261204
COMBINED_SEED = 20261205


def base_params(lbar):
    """The plain configuration's parameters for a mesh of mean edge lbar."""
    return dict(kappa_b=30.0, kappa_a=1e6, kappa_v=1e6, kappa_c=0.0, kappa_t=1e5, kappa_r=1e3,
                area_frac=1.0, volume_frac=0.8, curvature_frac=1.0,
                bond_lc0=1.25 * lbar, bond_lc1=0.75 * lbar, bond_a0=1.0, bond_r=2, bond_type=0,
                rep_lc1=1.5 * lbar, rep_r=2, rep_nlist=0, rep_rlist=1.5 * lbar, rep_exclusion=2,
                cont_delta=0.0, cont_lambda=1.0)


def single(level, seed=SINGLE_SEED):
    """(points, faces, mean edge, params) of one jittered icosphere."""
    v, f, lbar = meshes.jittered_icosphere(level, seed=seed, rel_sigma=0.01)
    return v, f, lbar, base_params(lbar)


def replicas(n_replicas=64, seed=REPLICA_SEED):
    """([(points, faces, params)] * n, ladder) of the batched-replica set."""
    rng = np.random.default_rng(seed)
    levels = rng.integers(3, 7, size=n_replicas)
    vfrac = rng.uniform(0.6, 0.95, size=n_replicas)
    ladder = np.sort(np.exp(rng.uniform(np.log(10.0), np.log(100.0), size=n_replicas)))
    out = []
    for r in range(n_replicas):
        v, f, lbar = meshes.jittered_icosphere(int(levels[r]), seed=seed + 1 + r, rel_sigma=0.01)
        out.append((v, f, dict(base_params(lbar), kappa_b=float(ladder[r]), volume_frac=float(vfrac[r]))))
    return out, ladder


def combined(n_replicas=32, seed=COMBINED_SEED, level=7):
    """([(points, faces, params)] * n, ladder) of the combined set: one
    icosphere, each replica jittered from its own seed."""
    v0, f = meshes.icosphere(level)
    f = np.ascontiguousarray(f.astype(np.int32))
    lbar = meshes.mean_edge_length(v0, f)
    rng = np.random.default_rng(seed)
    vfrac = rng.uniform(0.55, 0.9, size=n_replicas)
    ladder = np.sort(np.exp(rng.uniform(np.log(20.0), np.log(80.0), size=n_replicas)))
    out = []
    for r in range(n_replicas):
        v = np.ascontiguousarray(v0 + np.random.default_rng(seed + 1 + r).normal(0.0, 0.01 * lbar, v0.shape))
        out.append((v, f, dict(base_params(lbar), kappa_b=float(ladder[r]), volume_frac=float(vfrac[r]))))
    return out, ladder


def core_params(core, p):
    """trimem_amd.core.EnergyParams of a parameter dict."""
    e = core.EnergyParams()
    for k in ("kappa_b", "kappa_a", "kappa_v", "kappa_c", "kappa_t", "kappa_r",
              "area_frac", "volume_frac", "curvature_frac"):
        setattr(e, k, p.get(k, getattr(e, k)))
    b = core.BondParams()
    b.lc0, b.lc1, b.a0, b.r = p["bond_lc0"], p["bond_lc1"], p.get("bond_a0", 1.0), p.get("bond_r", 2)
    b.type = [core.BondType.Edge, core.BondType.Area, getattr(core.BondType, "None")][p.get("bond_type", 0)]
    e.bond_params = b
    r = core.SurfaceRepulsionParams()
    r.lc1, r.r, r.rlist = p["rep_lc1"], p.get("rep_r", 2), p["rep_rlist"]
    r.exclusion_level = p.get("rep_exclusion", 0)
    r.n_search = ["cell-list", "verlet-list"][p.get("rep_nlist", 0)]
    e.repulse_params = r
    c = core.ContinuationParams()
    c.delta, c.lam = p.get("cont_delta", 0.0), p.get("cont_lambda", 1.0)
    e.continuation_params = c
    return e
