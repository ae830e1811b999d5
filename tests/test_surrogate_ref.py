"""The CPU restatement of the surrogate network (tests/surrogate_ref.py)
against what the mathematics fixes, before it is used to check the GPU:

* autograd of the loss against central finite differences (float64, no
  bf16 rounding) along random directions of parameter space;
* rotation equivariance: rotating the mesh rotates the predicted forces;
* translation invariance;
* ring construction: every face (v, j_k, j_{k-1}) of a ring is a face of the
  mesh with the same orientation.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from trimem_amd import meshes as M
import surrogate_ref as S


def setup(level=1, seed=3):
    v, f, lbar = M.jittered_icosphere(level, seed)
    p = O.Params(kappa_b=1.0, kappa_a=10.0, kappa_v=10.0, kappa_c=1.0, kappa_t=1.0, volume_frac=0.9,
                 bond_lc0=1.3 * lbar, bond_lc1=0.7 * lbar)
    om = O.Mesh(v, f)
    oe = O.EnergyManager(om, p)
    rng = np.random.default_rng(seed)
    norm = {"mu": rng.normal(size=8) * 0.1, "sig": 1.0 + rng.uniform(size=8), "lbar": lbar}
    theta = rng.normal(size=S.OFF["n"]) * 0.2
    return v, f, om, oe, norm, theta


def features(v, f, om, oe, norm):
    return S.geometry(v, f, oe.vertex_properties(om), S.factors_of(oe, om), norm)


def test_rings_match_faces():
    v, f = M.icosphere(2)
    fs = {tuple(np.roll(t, -np.argmin(t))) for t in f}
    for vtx, r in enumerate(S.rings(f, len(v))):
        for k in range(len(r)):
            t = np.array([vtx, r[k], r[k - 1]])
            assert tuple(np.roll(t, -np.argmin(t))) in fs


def test_autograd_matches_finite_differences():
    v, f, om, oe, norm, theta = setup()
    fe, sp, nu = features(v, f, om, oe, norm)
    g = oe.gradient(om)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    L = S.loss(S.forward(th, fe, sp, nu, 2.0, exact=True), g, 2.0)
    (grad,) = torch.autograd.grad(L, th)
    rng = np.random.default_rng(0)
    for _ in range(4):
        d = torch.tensor(rng.normal(size=theta.shape), dtype=torch.float64)
        h = 1e-6
        with torch.no_grad():
            lp = S.loss(S.forward(th + h * d, fe, sp, nu, 2.0, exact=True), g, 2.0)
            lm = S.loss(S.forward(th - h * d, fe, sp, nu, 2.0, exact=True), g, 2.0)
        fd = float((lp - lm) / (2 * h))
        an = float(grad @ d)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-12) + 1e-10


def test_rotation_equivariance_and_translation_invariance():
    v, f, om, oe, norm, theta = setup(seed=5)
    th = torch.tensor(theta, dtype=torch.float64)
    F0 = S.forward(th, *features(v, f, om, oe, norm), 1.0, exact=True).numpy()
    q, _ = np.linalg.qr(np.random.default_rng(1).normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    for x, expect in ((v @ q.T, F0 @ q.T), (v + np.array([0.3, -1.2, 2.0]), F0)):
        om2 = O.Mesh(x, f)
        oe2 = O.EnergyManager(om2, oe_params(oe))
        F1 = S.forward(th, *features(x, f, om2, oe2, norm), 1.0, exact=True).numpy()
        assert np.allclose(F1, expect, rtol=0, atol=1e-9 * np.abs(F0).max())


def oe_params(oe):
    p = oe.params()
    return O.Params(**{k: getattr(p, k) for k, _ in p._fields_ if k != "pad_"})
