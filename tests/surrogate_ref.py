"""CPU restatement (torch) of the surrogate force network of
trimem_amd/csrc/tm_surrogate.hip, for the tests.

Features are built here from positions, the face list and the oracle's
per-vertex and global properties, independently of the GPU feature kernel.
``bf`` rounds matrix-product operands to bfloat16 as the GPU's MFMA operands
are rounded (round to nearest even), and ``exact=True`` drops that rounding
and runs in float64 (for finite-difference checks of the formulation).
"""
import numpy as np
import torch

W = 64
LAYERS = 3
FEAT = 8
FEATPAD = 16


def offsets():
    off = {"Win": 0, "bin": FEATPAD * W}
    base = off["bin"] + W
    for l in range(LAYERS):
        off[f"U{l}"] = base
        off[f"W{l}"] = base + W * W
        off[f"V{l}"] = base + 2 * W * W
        off[f"E{l}"] = base + 3 * W * W
        off[f"b{l}"] = base + 3 * W * W + 2 * W
        base += 3 * W * W + 3 * W
    off["ws"] = base
    off["bs"] = base + W
    off["wt"] = base + W + 1
    off["bt"] = base + 2 * W + 1
    off["n"] = base + 2 * W + 2
    return off


OFF = offsets()


def unpack(theta):
    t = theta
    p = {"Win": t[OFF["Win"]:OFF["Win"] + FEATPAD * W].reshape(FEATPAD, W), "bin": t[OFF["bin"]:OFF["bin"] + W]}
    for l in range(LAYERS):
        p[f"U{l}"] = t[OFF[f"U{l}"]:OFF[f"U{l}"] + W * W].reshape(W, W)
        p[f"W{l}"] = t[OFF[f"W{l}"]:OFF[f"W{l}"] + W * W].reshape(W, W)
        p[f"V{l}"] = t[OFF[f"V{l}"]:OFF[f"V{l}"] + W * W].reshape(W, W)
        p[f"E{l}"] = t[OFF[f"E{l}"]:OFF[f"E{l}"] + 2 * W].reshape(2, W)
        p[f"b{l}"] = t[OFF[f"b{l}"]:OFF[f"b{l}"] + W]
    p["ws"] = t[OFF["ws"]:OFF["ws"] + W]
    p["bs"] = t[OFF["bs"]]
    p["wt"] = t[OFF["wt"]:OFF["wt"] + W]
    p["bt"] = t[OFF["bt"]]
    return p


def rings(faces, n):
    """Neighbours of every vertex in the order j_0, j_1, ... such that each
    face (v, j_k, j_{k-1}) is one of v's counter-clockwise triangles (closed,
    consistently oriented mesh)."""
    nxt = [dict() for _ in range(n)]
    for a, b, c in faces:
        for v, p, q in ((a, b, c), (b, c, a), (c, a, b)):
            nxt[v][q] = p        # face (v, p, q): j_k = p follows j_{k-1} = q
    out = []
    for v in range(n):
        start = next(iter(nxt[v]))
        r = [start]
        while True:
            j = nxt[v][r[-1]]
            if j == start:
                break
            r.append(j)
        out.append(r)
    return out


def geometry(x, faces, vprops, factors, norm):
    """Input features f (N,16), spoke data and vertex normals (float64)."""
    n = len(x)
    rg = rings(faces, n)
    lbar = norm["lbar"]
    f = np.zeros((n, FEATPAD))
    spokes = []    # (v, j, eps0, eps1, ehat)
    nu = np.zeros((n, 3))
    for v in range(n):
        r = rg[v]
        k_ = len(r)
        e = x[r] - x[v]
        nsum = np.zeros(3)
        lsum = 0.0
        angles = 0.0
        for k in range(k_):
            nk = np.cross(e[k], e[k - 1])
            nk1 = np.cross(e[(k + 1) % k_], e[k])
            nsum += nk
            angles += np.arctan2(np.linalg.norm(nk), np.dot(e[k], e[k - 1]))
            l = np.linalg.norm(e[k])
            lsum += l
            s = np.dot(np.cross(nk, nk1), e[k]) / l
            c = np.dot(nk, nk1)
            spokes.append((v, r[k], (l - lbar) / lbar, np.arctan2(s, c), e[k] / l))
        nu[v] = nsum / np.linalg.norm(nsum)
        a, cv = vprops[v, 0], vprops[v, 2]
        raw = [a, cv / a, k_, lsum / k_, factors[0], factors[1], factors[2], 2 * np.pi - angles]
        f[v, :FEAT] = (np.array(raw) - norm["mu"]) / norm["sig"]
    return f, spokes, nu


def factors_of(oe, om):
    """Relative deviations A/A_ref - 1, V/V_ref - 1, C/C_ref - 1 of the oracle
    energy manager at its mesh (reference values as energy.cpp:30-45)."""
    P = oe.params()
    init = oe.initial_props()
    pr = oe.properties(om)
    lam = P.cont_lambda
    ref = [(1 - lam * (1 - P.area_frac)) * init[0], (1 - lam * (1 - P.volume_frac)) * init[1],
           (1 - lam * (1 - P.curvature_frac)) * init[2]]
    return [pr[i] / ref[i] - 1 for i in range(3)]


def bf(t):
    return t.to(torch.bfloat16).to(t.dtype)


def forward(theta, f, spokes, nu, sigma_F, exact=False):
    """Forces (N,3) as a torch tensor; theta a torch parameter vector."""
    dt = torch.float64 if exact else torch.float32
    rnd = (lambda t: t) if exact else bf
    p = unpack(theta.to(dt))
    silu = torch.nn.functional.silu
    f = torch.as_tensor(f, dtype=dt)
    vi = torch.tensor([s[0] for s in spokes])
    ji = torch.tensor([s[1] for s in spokes])
    eps = torch.tensor([[s[2], s[3]] for s in spokes], dtype=dt)
    ehat = torch.tensor(np.array([s[4] for s in spokes]), dtype=dt)
    nu = torch.as_tensor(nu, dtype=dt)
    n = f.shape[0]
    H = silu(rnd(f) @ rnd(p["Win"]) + p["bin"])
    for l in range(LAYERS):
        P = rnd(H) @ rnd(p[f"U{l}"])
        A = P[ji] + eps[:, 0:1] * p[f"E{l}"][0] + eps[:, 1:2] * p[f"E{l}"][1]
        M = torch.zeros(n, W, dtype=dt).index_add(0, vi, silu(A)) * (1.0 / 6.0)
        Z = rnd(torch.cat([H, M], 1)) @ rnd(torch.cat([p[f"W{l}"], p[f"V{l}"]], 0)) + p[f"b{l}"]
        H = H + silu(Z)
    s = (H[vi] * H[ji] * p["ws"]).sum(1) + p["bs"]
    t = (H * p["wt"]).sum(1) + p["bt"]
    F = torch.zeros(n, 3, dtype=dt).index_add(0, vi, s[:, None] * ehat) + t[:, None] * nu
    return sigma_F * F


def loss(F, grad_exact, sigma_F):
    """mean over vertices and components of (F / sigma_F + dE/dx / sigma_F)^2"""
    g = torch.as_tensor(grad_exact, dtype=F.dtype)
    r = F / sigma_F + g / sigma_F
    return (r * r).mean()
