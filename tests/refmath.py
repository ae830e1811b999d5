"""Independent FP64 formulations used to pin the oracle (test helper).

These are written from the mathematics of the Helfrich functional as trimem
defines it (src/kernel.h:110-137, src/mesh_properties.cpp:12-49), but with a
different structure from the oracle: sums run over faces and undirected edges
with numpy/torch vector ops instead of per-vertex half-edge circulation, and
gradients come from torch autograd instead of the hand-derived formulas of
src/mesh_util.h / src/mesh_properties.cpp:51-148.  A dropped term, a wrong
sign or a swapped vertex in either implementation shows up as a mismatch.
"""
import numpy as np
import torch


def edge_table(faces):
    """Undirected interior/boundary edges with their two opposite vertices.

    Returns (ij, k, l): ij[e] = (i, j) with the halfedge i->j belonging to a
    face (i, j, k); l is the vertex opposite in the face of j->i, or -1 on a
    boundary.
    """
    opp = {}
    for a, b, c in faces:
        opp[(a, b)] = c
        opp[(b, c)] = a
        opp[(c, a)] = b
    seen = set()
    ij, ks, ls = [], [], []
    for (a, b), c in opp.items():
        key = (min(a, b), max(a, b))
        if key in seen:
            continue
        seen.add(key)
        if (b, a) in opp:
            ij.append((a, b))
            ks.append(c)
            ls.append(opp[(b, a)])
        else:
            ij.append((a, b))
            ks.append(c)
            ls.append(-1)
    return np.array(ij, dtype=np.int64), np.array(ks, dtype=np.int64), np.array(ls, dtype=np.int64)


def one_ring(n, faces):
    nb = [set() for _ in range(n)]
    for a, b, c in faces:
        nb[a].update((b, c))
        nb[b].update((a, c))
        nb[c].update((a, b))
    return nb


def excluded_sets(n, faces, level):
    """Vertices excluded from i's repulsion at exclusion level 0/1/2, by BFS."""
    nb = one_ring(n, faces)
    out = []
    for i in range(n):
        s = {i}
        if level >= 1:
            s |= nb[i]
        if level >= 2:
            for j in nb[i]:
                s |= nb[j]
        out.append(s)
    return out


def brute_pairs(x, faces, dmax, level):
    """All ordered pairs (i, j) with |x_i - x_j| < dmax that are not excluded."""
    n = len(x)
    ex = excluded_sets(n, faces, level)
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    pairs = []
    for i in range(n):
        for j in np.nonzero(d[i] < dmax)[0]:
            if j not in ex[i]:
                pairs.append((i, int(j)))
    return np.array(pairs, dtype=np.int64).reshape((-1, 2))


def sector_normal(xa, xb, xc):
    """Normal (x_c - x_b) x (x_a - x_b) of the face (a, b, c), length 2*area."""
    return torch.cross(xc - xb, xa - xb, dim=-1)


def energy_terms(x, faces, params, pairs=None):
    """Global properties (A, V, C, B, T, R) of trimem's functional in torch.

    x: (N,3) float64 tensor (requires_grad for autograd), faces: (F,3) ints.
    params: dict with bond_type ('edge'/'area'), bond_lc0, bond_lc1, bond_a0,
    bond_r, rep_lc1, rep_r.  pairs: (P,2) ordered repulsion pairs.
    """
    f = torch.as_tensor(np.asarray(faces), dtype=torch.int64)
    n = x.shape[0]
    xa, xb, xc = x[f[:, 0]], x[f[:, 1]], x[f[:, 2]]
    nf = sector_normal(xa, xb, xc)
    Af = torch.linalg.norm(nf, dim=1) / 2
    Vf = (xa * torch.cross(xb, xc, dim=1)).sum(1) / 6

    ij, k, l = edge_table(np.asarray(faces))
    inner = l >= 0
    ij_t = torch.as_tensor(ij)
    xi, xj = x[ij_t[:, 0]], x[ij_t[:, 1]]
    le = torch.linalg.norm(xj - xi, dim=1)
    th = torch.zeros_like(le)
    if inner.any():
        ii = torch.as_tensor(np.nonzero(inner)[0])
        kk = torch.as_tensor(k[inner])
        ll = torch.as_tensor(l[inner])
        pi, pj = xi[ii], xj[ii]
        n0 = sector_normal(pi, pj, x[kk])          # face of i->j
        n1 = sector_normal(pj, pi, x[ll])          # face of j->i
        c = (n0 * n1).sum(1) / (torch.linalg.norm(n0, dim=1) * torch.linalg.norm(n1, dim=1))
        s = (torch.cross(n0, n1, dim=1) * (pj - pi)).sum(1)
        ang = torch.acos(torch.clamp(c, -1.0, 1.0))
        th = th.index_put((ii,), torch.where(s >= 0, ang, -ang))

    # per-vertex area and curvature
    a_v = torch.zeros(n, dtype=x.dtype).index_add(0, f.reshape(-1), Af.repeat_interleave(3) / 3)
    lt = le * th / 4
    c_v = torch.zeros(n, dtype=x.dtype).index_add(0, ij_t[:, 0], lt).index_add(0, ij_t[:, 1], lt)
    A = Af.sum()
    V = Vf.sum()
    C = (le * th).sum() / 2
    B = (2 * c_v * c_v / a_v).sum()

    if params.get("bond_type", "edge") == "edge":
        lc0, lc1, r = params["bond_lc0"], params["bond_lc1"], params["bond_r"]
        # both branches are disjoint for lc1 < lc0; torch.where evaluates both
        hi = torch.where(le > lc0, (r ** (r + 1)) * torch.clamp(le - lc0, min=0) ** r, torch.zeros_like(le))
        safe = torch.where(le < lc1, le, torch.full_like(le, 0.5 * lc1 if lc1 > 0 else 1.0))
        lo = torch.where(le < lc1, torch.exp(safe / (safe - lc1)) / safe ** r, torch.zeros_like(le))
        T = (hi + lo).sum()
    else:
        d = Af / params["bond_a0"] - 1.0
        T = (d * d).sum()

    R = torch.zeros((), dtype=x.dtype)
    if pairs is not None and len(pairs):
        p = torch.as_tensor(pairs)
        lr = torch.linalg.norm(x[p[:, 0]] - x[p[:, 1]], dim=1)
        lc1r, rr = params["rep_lc1"], params["rep_r"]
        R = (torch.exp(lr / (lr - lc1r)) / lr ** rr).sum()
    return A, V, C, B, T, R


def total_energy(props, ref, kappa):
    """trimem_energy (src/kernel.h:110-122) over torch scalars."""
    A, V, C, B, T, R = props
    A0, V0, C0 = ref
    kb, ka, kv, kc, kt, kr = kappa
    return (ka * (A / A0 - 1) ** 2 + kv * (V / V0 - 1) ** 2 + kc * (C / C0 - 1) ** 2
            + kt * T + kr * R + kb * B)
