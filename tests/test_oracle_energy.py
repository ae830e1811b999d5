"""Oracle energy, properties and gradient against known answers.

* closed forms for sphere and tube (tests/test_energy.py:15-33, rel 1e-2);
* the documented EnergyManager.energy value and gradient rows
  (doc/python-module.rst:60-97), reproduced to 1e-13 / print precision;
* an independent face/edge-centric torch formulation with autograd
  (tests/refmath.py), rel 1e-10, including repulsion pairs and both bond types;
* the reference's finite-difference test (tests/test_gradient.py:50-63);
* invariants: translation invariance of the A, C, B, T gradients and
  Euler homogeneity sum_v x_v . dV/dx_v = 3 V.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from trimem_amd import meshes as M
import refmath

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


def sphere(r, n=32):
    v, f = M.icosphere(int(np.log2(n)))
    return O.Mesh(v * r, f), (4 * np.pi * r ** 2, 4 / 3 * np.pi * r ** 3, 4 * np.pi * r, 8 * np.pi)


def tube(r, n=32):
    v, f = M.tube(length=1, radius=r, segments=n)
    return O.Mesh(v, f), (2 * np.pi * r, 2 / 3 * np.pi * r ** 2, np.pi, np.pi / r)


@pytest.mark.parametrize("kind,r", [("sphere", 1.0), ("sphere", 0.5), ("tube", 1.0), ("tube", 0.5)])
def test_closed_forms(kind, r):
    mesh, ref = (sphere if kind == "sphere" else tube)(r)
    em = O.EnergyManager(mesh, O.Params())
    p = em.properties(mesh)
    for k in range(4):
        assert abs(p[k] - ref[k]) / ref[k] < 1e-2


def doc_mesh():
    v, f = M.icosphere(3)
    return O.Mesh(v.astype(np.float32).astype(np.float64), f)


def doc_params():
    g = GOLD["doc_energy"]["params"]
    return O.Params(**g)


def test_doc_energy_value():
    m = doc_mesh()
    em = O.EnergyManager(m, doc_params())
    e = em.energy(m)
    ref = GOLD["doc_energy"]["energy"]
    assert abs(e - ref) / ref < 1e-13


def test_doc_gradient_rows():
    m = doc_mesh()
    em = O.EnergyManager(m, doc_params())
    g = em.gradient(m)
    rows = GOLD["doc_energy"]["gradient_rows_first"] + GOLD["doc_energy"]["gradient_rows_last"]
    for r in rows:
        r = np.array(r)
        d = np.linalg.norm(g - r, axis=1) / np.linalg.norm(r)
        assert d.min() < 5e-9
    # the STL reader numbers vertices by first appearance in the face list,
    # which puts icosphere vertices 0, 532, 196 in the first three rows
    first = GOLD["doc_energy"]["gradient_rows_first"]
    for vid, r in zip((0, 532, 196), first):
        assert np.linalg.norm(g[vid] - r) / np.linalg.norm(r) < 5e-9


def _torch_check(v, f, params, pairs_level=None, rel_e=1e-11, rel_g=1e-9):
    mesh = O.Mesh(v, f)
    em = O.EnergyManager(mesh, params)
    pairs = None
    if params.rep_lc1 > 0:
        pairs = refmath.brute_pairs(v, f, params.rep_lc1, params.rep_exclusion)
    tp = dict(bond_type="edge" if params.bond_type == 0 else "area", bond_lc0=params.bond_lc0,
              bond_lc1=params.bond_lc1, bond_a0=params.bond_a0, bond_r=params.bond_r,
              rep_lc1=params.rep_lc1, rep_r=params.rep_r)
    x0 = torch.tensor(v, dtype=torch.float64)
    with torch.no_grad():
        init = refmath.energy_terms(x0, f, tp, pairs)
    lam = params.cont_lambda
    ref = ((1 - lam * (1 - params.area_frac)) * init[0], (1 - lam * (1 - params.volume_frac)) * init[1],
           (1 - lam * (1 - params.curvature_frac)) * init[2])
    x = torch.tensor(v, dtype=torch.float64, requires_grad=True)
    props = refmath.energy_terms(x, f, tp, pairs)
    kap = (params.kappa_b, params.kappa_a, params.kappa_v, params.kappa_c, params.kappa_t, params.kappa_r)
    E = refmath.total_energy(props, ref, kap)
    (gt,) = torch.autograd.grad(E, x)
    po = em.properties(mesh)
    for k in range(6):
        tv = float(props[k])
        assert abs(po[k] - tv) <= rel_e * max(abs(tv), 1e-300) + 1e-14, (k, po[k], tv)
    eo = em.energy(mesh)
    assert abs(eo - float(E)) / abs(float(E)) < rel_e
    go = em.gradient(mesh)
    assert np.linalg.norm(go - gt.numpy()) / np.linalg.norm(gt.numpy()) < rel_g
    return em, mesh


def _params_l2(lbar, **kw):
    base = dict(kappa_b=30.0, kappa_a=1e4, kappa_v=1e4, kappa_c=1e3, kappa_t=1e2, kappa_r=1e1,
                area_frac=0.9, volume_frac=0.8, curvature_frac=1.1, bond_type=0,
                bond_lc0=1.02 * lbar, bond_lc1=0.98 * lbar, bond_a0=1.0, bond_r=2)
    base.update(kw)
    return O.Params(**base)


@pytest.mark.parametrize("bond", [0, 1])
def test_energy_gradient_vs_torch(bond):
    v, f, lbar = M.jittered_icosphere(2, seed=7, rel_sigma=0.05)
    a = 4 * np.pi / len(f)
    _torch_check(v, f, _params_l2(lbar, bond_type=bond, bond_a0=a))


@pytest.mark.parametrize("nlist,excl", [(0, 0), (0, 1), (0, 2), (1, 2)])
def test_repulsion_vs_torch(nlist, excl):
    """Flattened sphere: top and bottom sheets come within lc1 of each other."""
    v, f, lbar = M.jittered_icosphere(3, seed=3, rel_sigma=0.02)
    v = v * np.array([1.0, 1.0, 0.06])
    lc1 = 2.0 * lbar
    p = _params_l2(lbar, rep_lc1=lc1, rep_rlist=lc1, rep_nlist=nlist, rep_exclusion=excl, rep_r=2,
                   kappa_r=1.0)
    pairs = refmath.brute_pairs(v, f, lc1, excl)
    assert len(pairs) > 50
    _torch_check(v, f, p)


@pytest.mark.parametrize("bond", [0, 1])
def test_fd_gradient(bond):
    """tests/test_gradient.py:14-63: kappas 1, fracs 0.8, lc0/lc1 = 1.15/0.85 l."""
    v, f = M.icosphere(3)
    m = O.Mesh(v, f)
    a, l = m.avg_tri_props()
    p = O.Params(kappa_b=1.0, kappa_a=1.0, kappa_v=1.0, kappa_c=1.0, kappa_t=1.0,
                 area_frac=0.8, volume_frac=0.8, curvature_frac=0.8, bond_type=bond,
                 bond_lc0=1.15 * l, bond_lc1=0.85 * l, bond_a0=a)
    em = O.EnergyManager(m, p)
    ref = em.fd_gradient(m, 1e-8)
    g = em.gradient(m)
    assert np.linalg.norm(g - ref) / np.linalg.norm(ref) < 1e-4


def test_gradient_invariants():
    v, f, lbar = M.jittered_icosphere(3, seed=11, rel_sigma=0.05)
    m = O.Mesh(v, f)
    # isolate single terms by switching the others off; fracs put each penalty
    # away from its minimum so its gradient is not zero
    def grad(**kw):
        base = dict(kappa_b=0.0, kappa_a=0.0, kappa_v=0.0, kappa_c=0.0, kappa_t=0.0,
                    area_frac=0.9, volume_frac=0.9, curvature_frac=0.9,
                    bond_lc0=0.9 * lbar, bond_lc1=0.5 * lbar)
        base.update(kw)
        em = O.EnergyManager(m, O.Params(**base))
        return em, em.gradient(m)
    for k in ("kappa_a", "kappa_c", "kappa_b", "kappa_t"):
        _, g = grad(**{k: 1.0})
        assert np.linalg.norm(g.sum(0)) < 1e-10 * np.abs(g).sum(), k
    em, g = grad(kappa_v=1.0)
    vol = em.properties(m)[1]
    ref_v = 0.9 * em.initial_props()[1]
    fac = 2.0 / ref_v * (vol / ref_v - 1.0)
    assert abs((g * v).sum() / fac - 3 * vol) < 1e-10 * abs(3 * vol)
