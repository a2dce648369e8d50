"""The FP64 CPU tape (tests/tape.py) against the oracle, and its gradients
against central finite differences.  CPU only."""
import numpy as np
import pytest
import torch

import mpmtest as mt
import tape

sq2 = lambda X: (np.abs(X[:, 0] - 0.5) < 0.2) & (np.abs(X[:, 1] - 0.6) < 0.2)
sph3 = lambda X: ((X - np.array([0.5, 0.5, 0.5])) ** 2).sum(1) < 0.2 ** 2


def scene(dim, model, wtype, seed=1234):
    N = 16 if dim == 2 else 8
    walls = mt.wall_boxes(dim, [N] * dim, [0] * dim, [1] * dim, flags=(1 << (2 * dim)) - 1, inset=0.3,
                          type_=wtype, friction=0.0)
    return mt.make_scene(dim, N, prec=8, model=model, shape=sq2 if dim == 2 else sph3, v0=[0.5, -2.0, 0.3][:dim],
                         vrand=1.0, brand=1e-2, frand=0.1, walls=walls, dt=1e-3, per_particle_E=True, seed=seed)


CASES = [(2, mt.COROTATED, mt.STICK), (2, mt.NEO_HOOKEAN, mt.SLIP), (3, mt.COROTATED, mt.SEPARATE),
         (3, mt.NEO_HOOKEAN, mt.STICK)]


@pytest.mark.parametrize("dim,model,wtype", CASES)
def test_tape_matches_oracle(dim, model, wtype):
    sc = scene(dim, model, wtype)
    steps = 6
    orc = mt.oracle_run(sc, steps, prec=8, want_grid=False)
    assert orc["valid"].all()
    st, prm = tape.to_state(sc)
    with torch.no_grad():
        X, V, F, B = tape.run(st, prm, sc, steps)
    for name, a, b in (("X", X, orc["X"]), ("V", V, orc["V"]), ("F", tape.flat(F), orc["F"]),
                       ("B", tape.flat(B), orc["B"])):
        a = a.numpy().reshape(b.shape)
        err = np.abs(a - b).max() / np.abs(b).max()
        print(dim, model, wtype, name, f"{err:.2e}")
        assert err < 1e-11, name


@pytest.mark.parametrize("dim,model,wtype", CASES)
def test_tape_gradients_match_finite_differences(dim, model, wtype):
    sc = scene(dim, model, wtype)
    steps = 4
    rng = np.random.default_rng(3)
    st, prm = tape.to_state(sc)
    n = st[0].shape[0]
    c = torch.tensor(rng.standard_normal((n, dim)))

    def loss(V0, mu, la):
        s = [st[0], V0, st[2], st[3]]
        X, V, F, B = tape.run(s, [prm[0], prm[1], mu, la], sc, steps)
        return (c * X).sum() + 0.1 * (c * V).sum() + 1e-3 * (F ** 2).sum() + 1e-4 * (B ** 2).sum()

    V0 = st[1].clone().requires_grad_(True)
    mu = prm[2].clone().requires_grad_(True)
    la = prm[3].clone().requires_grad_(True)
    L = loss(V0, mu, la)
    gV, gmu, gla = torch.autograd.grad(L, (V0, mu, la))
    for _ in range(3):
        dV = torch.tensor(rng.standard_normal(V0.shape))
        dm = torch.tensor(rng.standard_normal(n)) * mu.detach() * 1e-2
        dl = torch.tensor(rng.standard_normal(n)) * la.detach() * 1e-2
        eps = 1e-5
        with torch.no_grad():
            lp = loss(V0 + eps * dV, mu + eps * dm, la + eps * dl)
            lm = loss(V0 - eps * dV, mu - eps * dm, la - eps * dl)
        fd = float((lp - lm) / (2 * eps))
        an = float((gV * dV).sum() + (gmu * dm).sum() + (gla * dl).sum())
        print(dim, model, wtype, "fd", fd, "autograd", an)
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-8)
