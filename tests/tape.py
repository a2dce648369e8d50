"""FP64 CPU tape of the explicit APIC MPM step: a differentiable PyTorch
restatement used as the reference for the adjoint kernels (test
infrastructure only).

It computes the same step as the oracle (MPM_DRIVER::Advance_One_Time_Step,
Hybrid_Methods/Examples_And_Drivers/MPM_DRIVER.cpp:141-157, symplectic Euler,
quadratic B-splines, APIC, flip 0) in the algebraically equal forms the GPU
uses, so that torch.autograd yields the exact vector-Jacobian products of one
step:
  * P2G with the elastic impulse folded into the momentum,
    p_i = sum_p w m (v + B D^-1 (x_i - x_p)) - dt V0 P(F) F^T grad w
    (MPM_DRIVER.cpp:245-306, Forces/MPM_FINITE_ELEMENTS.cpp:84-99);
  * v_i = p_i / m_i + dt g, then box collision objects in order
    (System/MPM_OBJECTIVE.cpp:150-182); no friction;
  * G2P: v_p = sum w v_i, B = sum w v_i (x_i - x_p)^T, F <- (I + dt grad v) F,
    x <- x + dt v_p (MPM_DRIVER.cpp:450-515).
Fixed corotated uses the polar factor from Newton's iteration
R <- (R + R^-T) / 2 (equal to U V^T of the SVD for det F > 0, and smooth where
singular values repeat); neo-Hookean uses tau = V0 (mu (F F^T - I) + lambda ln J I)
(the branch J >= failure threshold).  Particles are assumed to stay inside
the domain (no invalidation).  Checked against the oracle in
tests/test_tape.py and against central finite differences.
"""
import itertools

import numpy as np
import torch

FLT_MAX = float(np.finfo(np.float32).max)


def to_state(sc, requires_grad=False):
    d = sc["dim"]
    t = lambda a: torch.tensor(np.asarray(a, np.float64))
    X, V = t(sc["X"]).reshape(-1, d), t(sc["V"]).reshape(-1, d)
    F = t(sc["F"]).reshape(-1, d, d).transpose(1, 2).contiguous()  # column-major -> F[p][i][j]
    B = t(sc["B"]).reshape(-1, d, d).transpose(1, 2).contiguous()
    st = [X, V, F, B]
    if requires_grad:
        for a in st:
            a.requires_grad_(True)
    prm = [t(sc["mass"]), t(sc["volume"]), t(sc["mu"]), t(sc["lambda"])]
    return st, prm


def flat(M):
    """(n, d, d) matrices -> (n, d*d) column-major rows."""
    return M.transpose(1, 2).reshape(M.shape[0], -1)


def polar_rotation(F, iters=14):
    R = F
    for _ in range(iters):
        R = 0.5 * (R + torch.linalg.inv(R).transpose(1, 2))
    return R


def kirchhoff(F, vol, mu, la, model):
    d = F.shape[-1]
    I = torch.eye(d, dtype=F.dtype)
    J = torch.linalg.det(F)
    if model == 0:
        R = polar_rotation(F)
        tau = 2 * mu[:, None, None] * (F - R) @ F.transpose(1, 2) + (la * (J - 1) * J)[:, None, None] * I
    else:
        tau = mu[:, None, None] * (F @ F.transpose(1, 2) - I) + (la * torch.log(J))[:, None, None] * I
    return vol[:, None, None] * tau


def box_phi(lo, hi, X):
    c = 0.5 * (lo + hi)
    hl = 0.5 * (hi - lo)
    phi = (X - c).abs() - hl
    inside = (phi <= 0).all(dim=1)
    out = torch.sqrt((phi.clamp(min=0) ** 2).sum(1))
    return torch.where(inside, phi.max(dim=1).values, out)


def box_normal(lo, hi, X):
    # Core/Math_Tools/RANGE.cpp:38-63; only the inside branch reaches a collision
    pmax = X - hi
    pmin = lo - X
    m = torch.maximum(pmin, pmax)
    axis = m.argmax(dim=1)
    n = torch.zeros_like(X)
    sgn = torch.where(pmax.gather(1, axis[:, None]) > pmin.gather(1, axis[:, None]), 1.0, -1.0).to(X.dtype)
    n.scatter_(1, axis[:, None], sgn)
    return n


class Grid:
    def __init__(self, sc):
        self.d = d = sc["dim"]
        self.N = np.array(sc["N"], np.int64).reshape(-1)
        if self.N.size == 1:
            self.N = np.repeat(self.N, d)
        self.lo = torch.tensor(np.asarray(sc["lo"], np.float64)[:d])
        self.hi = torch.tensor(np.asarray(sc["hi"], np.float64)[:d])
        self.h = (self.hi - self.lo) / torch.tensor(self.N, dtype=torch.float64)
        self.G = self.N + 6
        self.nn = int(np.prod(self.G))
        self.offsets = torch.tensor(list(itertools.product(range(3), repeat=d)), dtype=torch.int64)  # (K, d)
        stride = np.ones(d, np.int64)
        for a in range(d - 2, -1, -1):
            stride[a] = stride[a + 1] * self.G[a + 1]
        self.stride = torch.tensor(stride)

    def node_centers(self, idx):
        """flat node indices -> positions (ghost 3, cell-centred)."""
        r = idx.clone()
        coords = []
        for a in range(self.d):
            coords.append(r // self.stride[a])
            r = r % self.stride[a]
        c = torch.stack(coords, 1).to(torch.float64)
        return self.lo + (c - 3 + 0.5) * self.h


def step(state, params, sc, grid=None):
    X, V, F, B = state
    m, vol, mu, la = params
    gr = grid or Grid(sc)
    d, h, lo = gr.d, gr.h, gr.lo
    dt = float(sc["dt"])
    g = torch.tensor(np.asarray(sc["gravity"], np.float64)[:d])
    inv_h = 1.0 / h
    base = torch.floor(((X.detach() - h) - lo) * inv_h)
    cen = lo + (base + 0.5) * h
    xe = X - cen
    xx = xe * inv_h
    q0, q1, q2 = xx - 1.5, xx - 1.0, xx - 0.5
    w1d = torch.stack([0.5 * q0 * q0, 0.75 - q1 * q1, 0.5 * q2 * q2], -1)  # (n, d, 3)
    dw1d = torch.stack([q0 * inv_h, -2.0 * q1 * inv_h, q2 * inv_h], -1)
    o = gr.offsets  # (K, d)
    K = o.shape[0]
    wsel = torch.stack([w1d[:, a, o[:, a]] for a in range(d)], -1)   # (n, K, d)
    dwsel = torch.stack([dw1d[:, a, o[:, a]] for a in range(d)], -1)
    wk = wsel.prod(-1)                                                 # (n, K)
    gk = torch.stack([dwsel[..., c] * torch.stack([wsel[..., a] for a in range(d) if a != c], -1).prod(-1)
                      for c in range(d)], -1)                          # (n, K, d)
    dz = o.to(torch.float64) * h - xe[:, None, :]                      # (n, K, d)
    node = ((base.to(torch.int64)[:, None, :] + 3 + o) * gr.stride).sum(-1)  # (n, K)
    Dinv = 4.0 * inv_h * inv_h
    tau = kirchhoff(F, vol, mu, la, sc["model"])
    mom = (m[:, None, None] * wk[..., None] * (V[:, None, :] + torch.einsum("pij,pkj->pki", B, Dinv * dz))
           - dt * torch.einsum("pij,pkj->pki", tau, gk))
    nf = node.reshape(-1)
    gm = torch.zeros(gr.nn, dtype=torch.float64).index_add(0, nf, (m[:, None] * wk).reshape(-1))
    gp = torch.zeros(gr.nn, d, dtype=torch.float64).index_add(0, nf, mom.reshape(-1, d))
    act = torch.nonzero(gm.detach() > 0).reshape(-1)
    v = gp[act] / gm[act][:, None] + dt * g
    if sc.get("walls"):
        X0 = gr.node_centers(act)
        for w in sc["walls"]:
            assert w["friction"] == 0 or w["type"] == 0, "the tape has no friction"
            wlo = torch.tensor(np.asarray(w["lo"], np.float64)[:d])
            whi = torch.tensor(np.asarray(w["hi"], np.float64)[:d])
            with torch.no_grad():
                phi0 = box_phi(wlo, whi, X0)
                Xd = X0 + dt * v
                phi_d = box_phi(wlo, whi, Xd) - phi0.clamp(max=0)
                if w["type"] == 2:
                    hit = phi_d <= 1e-15
                else:
                    hit = phi0 <= 0
                if w["type"] != 0:
                    hit = hit & (phi_d <= FLT_MAX)
                nrm = box_normal(wlo, whi, Xd)
            if not bool(hit.any()):
                continue
            if w["type"] == 0:
                v = torch.where(hit[:, None], torch.zeros_like(v), v)
            else:
                phi = box_phi(wlo, whi, X0 + dt * v) - phi0.clamp(max=0)
                vn = v - (phi / dt)[:, None] * nrm
                v = torch.where(hit[:, None], vn, v)
    vg = torch.zeros(gr.nn, d, dtype=torch.float64).index_copy(0, act, v)
    vk = vg[node]                                                     # (n, K, d)
    vp = (wk[..., None] * vk).sum(1)
    Gm = torch.einsum("pkc,pkj->pcj", vk, gk)
    Bn = torch.einsum("pk,pkc,pkj->pcj", wk, vk, dz)
    Fn = F + dt * Gm @ F
    Xn = X + dt * vp
    return [Xn, vp, Fn, Bn]


def run(state, params, sc, nsteps):
    gr = Grid(sc)
    for _ in range(nsteps):
        state = step(state, params, sc, gr)
    return state
