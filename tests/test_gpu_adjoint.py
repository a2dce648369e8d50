"""Adjoint of the HIP step (mpmx_tape_begin / mpmx_backward) against the FP64
CPU tape (tests/tape.py: torch.autograd over the same step) and central
finite differences.

FP64 adjoints match the tape to 1e-9 relative (max over each quantity) after
one step and to 1e-8 over seven checkpointed steps.  The FP32 adjoint (stress
adjoint in FP32, FP64 only for inverted or nearly singular F) is compared
with the FP64 tape of the same float inputs to 1e-4.  The finite-difference
checks: the 2D variant of the differentiable configuration (32^2 grid, ~1k
particles, 200 steps, loss |mean x(T) - target|^2) and the 3D configuration
at its size (64^3, 56,581 particles, 200 steps, a checkpoint every 20, FP32
adjoint), central differences (eps 1e-4) of the GPU FP64 forward along random
directions of the initial velocities and of log E, relative error <= 1e-3.
"""
import numpy as np
import pytest
import torch

import mpmtest as mt
import tape
from mpm_autograd import MPMRollout, Rollout
from physbam_amd import mpmx

pytestmark = pytest.mark.gpu

sq2 = lambda X: (np.abs(X[:, 0] - 0.5) < 0.2) & (np.abs(X[:, 1] - 0.6) < 0.2)
sph3 = lambda X: ((X - np.array([0.5, 0.5, 0.5])) ** 2).sum(1) < 0.25 ** 2


def scene(dim, prec, model, wtype, seed=1234):
    N = 32 if dim == 2 else 16
    walls = mt.wall_boxes(dim, [N] * dim, [0] * dim, [1] * dim, flags=(1 << (2 * dim)) - 1, inset=0.28,
                          type_=wtype, friction=0.0)
    return mt.make_scene(dim, N, prec=prec, model=model, shape=sq2 if dim == 2 else sph3,
                         v0=[0.5, -2.0, 0.3][:dim], vrand=1.0, brand=1e-2, frand=0.1, walls=walls, dt=1e-3,
                         per_particle_E=True, seed=seed)


def gpu_backward(sc, steps, k, adj, pre_steps=0):
    d = sc["dim"]
    with mpmx.Context(d, sc["N"], sc["lo"], sc["hi"], precision=sc["prec"]) as ctx:
        ctx.set_constitutive(sc["model"], 0.25)
        ctx.set_gravity(sc["gravity"])
        for w in sc["walls"]:
            ctx.add_collision_box(w["type"], w["friction"], w["lo"][:d], w["hi"][:d])
        load = lambda: ctx.set_particles(sc["X"], sc["V"], sc["F"], sc["B"], sc["mass"], sc["volume"], sc["mu"],
                                         sc["lambda"], sc["valid"])
        if pre_steps:  # untaped steps first, then start again from the same state
            load()
            ctx.step(sc["dt"], pre_steps)
        load()
        ctx.tape_begin(k)
        ctx.step(sc["dt"], steps)
        fin = ctx.get_particles()
        g = ctx.backward(*adj)
        after = ctx.get_particles()
    assert np.array_equal(fin["X"], after["X"])  # the state is left at the final step
    return g


def tape_backward(sc, steps, adj):
    st, prm = tape.to_state(sc, requires_grad=True)
    mu = prm[2].clone().requires_grad_(True)
    la = prm[3].clone().requires_grad_(True)
    X, V, F, B = tape.run(st, [prm[0], prm[1], mu, la], sc, steps)
    d = sc["dim"]
    n = X.shape[0]
    t = lambda a, shape: torch.tensor(np.asarray(a, np.float64)).reshape(shape)
    L = ((X * t(adj[0], (n, d))).sum() + (V * t(adj[1], (n, d))).sum()
         + (tape.flat(F) * t(adj[2], (n, d * d))).sum() + (tape.flat(B) * t(adj[3], (n, d * d))).sum())
    gX, gV, gF, gB, gmu, gla = torch.autograd.grad(L, st + [mu, la])
    return dict(X=gX.numpy(), V=gV.numpy(), F=tape.flat(gF).numpy(), B=tape.flat(gB).numpy(), mu=gmu.numpy(),
                **{"lambda": gla.numpy()})


def random_adjoint(sc, seed):
    rng = np.random.default_rng(seed)
    n, d = sc["X"].shape[0], sc["dim"]
    return [rng.standard_normal((n, d)), rng.standard_normal((n, d)), rng.standard_normal((n, d * d)),
            1e-2 * rng.standard_normal((n, d * d))]


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


CASES = [(2, mt.COROTATED, mt.STICK), (2, mt.NEO_HOOKEAN, mt.SLIP), (3, mt.COROTATED, mt.SEPARATE),
         (3, mt.NEO_HOOKEAN, mt.SLIP), (3, mt.COROTATED, mt.STICK)]


@pytest.mark.parametrize("dim,model,wtype", CASES)
@pytest.mark.parametrize("steps,k", [(1, 1), (7, 3)])
def test_adjoint_matches_tape_fp64(dim, model, wtype, steps, k):
    sc = scene(dim, 8, model, wtype)
    adj = random_adjoint(sc, 5)
    tag = f"adjoint d{dim} m{model} w{wtype} steps{steps} k{k} n={sc['X'].shape[0]}"
    print(tag, flush=True)
    g = gpu_backward(sc, steps, k, adj)
    r = tape_backward(sc, steps, adj)
    errs = {q: rel(g[q], r[q]) for q in ("X", "V", "F", "B", "mu", "lambda")}
    print(tag, {q: f"{e:.1e}" for q, e in errs.items()}, flush=True)
    tol = 1e-9 if steps == 1 else 1e-8
    for q, e in errs.items():
        assert e <= tol, (q, e)


def test_adjoint_after_untaped_steps_on_the_same_context():
    """Untaped steps before mpmx_tape_begin (an odd number, so the grid arrays
    and tile lists of the step pipeline have swapped) leave the reverse pass
    unchanged: it matches a fresh context to round-off."""
    sc = scene(3, 8, mt.COROTATED, mt.SLIP)
    adj = random_adjoint(sc, 5)
    ref = gpu_backward(sc, 7, 3, adj)
    got = gpu_backward(sc, 7, 3, adj, pre_steps=3)
    errs = {q: rel(got[q], ref[q]) for q in ("X", "V", "F", "B", "mu", "lambda")}
    print("adjoint after untaped steps", {q: f"{e:.1e}" for q, e in errs.items()}, flush=True)
    for q, e in errs.items():
        assert e <= 1e-9, (q, e)


@pytest.mark.parametrize("dim,model,wtype", CASES[:3])
def test_adjoint_fp32_against_fp64_tape(dim, model, wtype):
    sc = scene(dim, 4, model, wtype)
    adj = [a.astype(np.float32).astype(np.float64) for a in random_adjoint(sc, 9)]
    g = gpu_backward(sc, 5, 2, adj)
    sc64 = dict(sc)
    for q in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda"):
        sc64[q] = np.asarray(sc[q], np.float64)
    r = tape_backward(sc64, 5, adj)
    errs = {q: rel(g[q], r[q]) for q in ("X", "V", "F", "B", "mu", "lambda")}
    print("fp32 adjoint", dim, model, wtype, {q: f"{e:.1e}" for q, e in errs.items()}, flush=True)
    for q, e in errs.items():
        assert e <= 1e-4, (q, e)


def test_rollout_gradients_match_finite_differences():
    nu = 0.3
    sc = mt.make_scene(2, 32, prec=8, model=mt.COROTATED,
                       shape=lambda X: (np.abs(X[:, 0] - 0.45) < 0.13) & (np.abs(X[:, 1] - 0.55) < 0.13),
                       v0=[0.4, 0.2], vrand=0.5, walls=[], dt=1e-4)
    n = sc["X"].shape[0]
    steps, k = 200, 20
    print("fd rollout particles", n, flush=True)
    rng = np.random.default_rng(2)
    logE = torch.tensor(rng.uniform(np.log(1e3), np.log(1e4), n), requires_grad=True)
    V0 = torch.tensor(sc["V"], requires_grad=True)
    target = torch.tensor([0.7, 0.3], dtype=torch.float64)
    sim = Rollout(sc, steps, k)

    def loss(V0, logE):
        E = torch.exp(logE)
        mu = E / (2 * (1 + nu))
        lam = E * nu / ((1 + nu) * (1 - 2 * nu))
        X = MPMRollout.apply(V0, mu, lam, sim)
        return ((X.mean(0) - target) ** 2).sum()

    try:
        L = loss(V0, logE)
        L.backward()
        for _ in range(3):
            dv = torch.tensor(rng.standard_normal(V0.shape))
            de = torch.tensor(rng.standard_normal(n))
            an = float((V0.grad * dv).sum() + (logE.grad * de).sum())
            eps = 1e-4
            with torch.no_grad():
                lp = float(loss(V0 + eps * dv, logE + eps * de))
                lm = float(loss(V0 - eps * dv, logE - eps * de))
            fd = (lp - lm) / (2 * eps)
            print("fd", fd, "adjoint", an, flush=True)
            assert abs(fd - an) <= 1e-3 * abs(fd)
    finally:
        sim.close()


def test_reverse_passes_in_a_row_on_one_context():
    """Tapes of different scenes on one context (other particle counts and Lame
    parameters, as in a training loop): the replay engine of the overlapped
    reverse pass is re-sized and re-loaded for each, so every reverse pass
    matches a fresh context's to round-off."""
    scs = [scene(3, 8, mt.COROTATED, mt.SLIP, seed=s) for s in (1234, 1234, 99)]
    scs[1]["mu"] = scs[1]["mu"] * 1.7  # the same count with other Lame parameters, then another count
    adjs = [random_adjoint(sc, 5 + i) for i, sc in enumerate(scs)]
    sc0 = scs[0]
    with mpmx.Context(3, sc0["N"], sc0["lo"], sc0["hi"], precision=8) as ctx:
        ctx.set_constitutive(sc0["model"], 0.25)
        ctx.set_gravity(sc0["gravity"])
        for w in sc0["walls"]:
            ctx.add_collision_box(w["type"], w["friction"], w["lo"], w["hi"])
        for i, (sc, adj) in enumerate(zip(scs, adjs)):
            print(f"reverse pass {i}: n={sc['X'].shape[0]}", flush=True)
            ctx.set_particles(sc["X"], sc["V"], sc["F"], sc["B"], sc["mass"], sc["volume"], sc["mu"], sc["lambda"],
                              sc["valid"])
            ctx.tape_begin(3)
            ctx.step(sc["dt"], 7)
            got = ctx.backward(*adj)
            ref = gpu_backward(sc, 7, 3, adj)
            errs = {q: rel(got[q], ref[q]) for q in ("X", "V", "F", "B", "mu", "lambda")}
            print(f"  {', '.join(f'{q} {e:.1e}' for q, e in errs.items())}", flush=True)
            for q, e in errs.items():
                assert e <= 1e-9, (i, q, e)


def test_scene_cannot_change_while_taping():
    """The reverse pass replays the recorded steps from the scene as it then is,
    so while a tape is open every call that would change the particles, the
    model, gravity or the collision objects fails with MPMX_ESTATE (on the
    host, before any kernel runs), and the tape's gradients are those of an
    undisturbed run."""
    sc = scene(2, 8, mt.COROTATED, mt.SLIP, seed=7)
    d = sc["dim"]
    n = sc["X"].shape[0]
    rng = np.random.default_rng(3)
    adj = (rng.normal(size=(n, d)), None, None, None)
    want = gpu_backward(sc, 4, 2, adj)
    with mpmx.Context(d, sc["N"], sc["lo"], sc["hi"], precision=8) as ctx:
        ctx.set_constitutive(sc["model"], 0.25)
        ctx.set_gravity(sc["gravity"])
        for w in sc["walls"]:
            ctx.add_collision_box(w["type"], w["friction"], w["lo"][:d], w["hi"][:d])
        args = (sc["X"], sc["V"], sc["F"], sc["B"], sc["mass"], sc["volume"], sc["mu"], sc["lambda"], sc["valid"])
        ctx.set_particles(*args)
        ctx.tape_begin(2)
        ctx.step(sc["dt"], 2)
        w0 = sc["walls"][0]
        calls = [lambda: ctx.set_particles(*args),
                 lambda: ctx.set_particles(*(a[: n // 2] for a in args)),
                 lambda: ctx.set_gravity([0.0, -1.0]),
                 lambda: ctx.set_constitutive(mt.NEO_HOOKEAN, 0.25),
                 lambda: ctx.add_collision_box(w0["type"], 0.0, w0["lo"][:d], w0["hi"][:d]),
                 lambda: ctx.clear_collision_objects()]
        for call in calls:
            with pytest.raises(mpmx.MpmxError) as e:
                call()
            assert e.value.code == -3  # MPMX_ESTATE
        ctx.step(sc["dt"], 2)
        got = ctx.backward(*adj)
        # the tape is consumed: the scene can change again
        ctx.set_gravity(sc["gravity"])
    for k in ("X", "V", "mu", "lambda"):
        np.testing.assert_allclose(got[k], want[k], rtol=0, atol=1e-12 * max(np.abs(want[k]).max(), 1e-30))


def test_differentiable_workload_gradients_match_fp64_finite_differences():
    """The 3D differentiable workload at its size (64^3 grid, one 0.3^3 box at
    8 ppc: 56,581 particles, fixed corotated E 1e4, nu 0.3, gravity, dt 1e-4,
    200 steps, a checkpoint every 20), FP32 adjoint, with E per particle
    log-uniform in [1e3, 1e5].  Each particle starts with a random velocity
    (U[-1, 1]^3 m/s) so the box deforms and its elastic waves make the loss
    depend on E; the loss is the squared distance of the mean final position
    of one octant of the box (x0, y0, z0 > 0.5) to (0.7, 0.3, 0.5) -- the
    mean over all particles would follow the total momentum alone and not
    depend on E.  The adjoint's directional derivatives
    along a random direction of the initial velocities and, separately, of
    log E per particle match central differences (eps 1e-4) of the GPU FP64
    forward to 1e-3 relative."""
    cells, dt, steps, k = 64, 1e-4, 200, 20
    box = [dict(type=mpmx.SHAPE_BOX, a=[0.35] * 3, b=[0.65] * 3, density=1000.0, E=1e4, nu=0.3, v=[0.0] * 3)]
    P = mpmx.seed_particles(3, cells, [0.0] * 3, [1.0] * 3, box, ppc_axis=2, seed=1234, precision=mpmx.FP32)
    n = P["X"].shape[0]
    assert n == 56581
    rng = np.random.default_rng(5)
    V0 = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    nu = 0.3
    logE = rng.uniform(np.log(1e3), np.log(1e5), n)
    lame = lambda lE: (np.exp(lE) / (2 * (1 + nu)), np.exp(lE) * nu / ((1 + nu) * (1 - 2 * nu)))
    upper = np.all(P["X"] > 0.5, axis=1)
    target = np.array([0.7, 0.3, 0.5])
    nu_ = int(upper.sum())

    def context(prec):
        ctx = mpmx.Context(3, cells, [0.0] * 3, [1.0] * 3, precision=prec)
        ctx.set_constitutive(mpmx.FIXED_COROTATED)
        ctx.set_gravity([0.0, -9.8, 0.0])
        return ctx

    def load(ctx, V, lE, dt_np):
        mu, la = lame(lE)
        c = lambda a: np.asarray(a, dt_np)
        ctx.set_particles(c(P["X"]), c(V), c(P["F"]), c(P["B"]), c(P["mass"]), c(P["volume"]), c(mu), c(la),
                          P["valid"])

    # FP32 adjoint of the taped rollout
    with context(mpmx.FP32) as ctx:
        load(ctx, V0, logE, np.float32)
        ctx.tape_begin(k)
        ctx.step(dt, steps)
        X = ctx.get_particles()["X"].astype(np.float64)
        mean = X[upper].mean(0)
        dX = np.zeros((n, 3))
        dX[upper] = 2 * (mean - target) / nu_
        g = ctx.backward(dX=dX)
    mu, la = lame(logE)
    g_logE = g["mu"] * mu + g["lambda"] * la  # mu and lambda are proportional to E

    def loss64(V, lE):
        with context(mpmx.FP64) as c64:
            load(c64, V, lE, np.float64)
            c64.step(dt, steps)
            Xf = c64.get_particles()["X"]
        return float(((Xf[upper].mean(0) - target) ** 2).sum())

    eps = 1e-4
    V64 = V0.astype(np.float64)
    dv = rng.standard_normal((n, 3))
    de = rng.standard_normal(n)
    for name, an, plus, minus in (
            ("v(0)", float((g["V"] * dv).sum()), lambda: loss64(V64 + eps * dv, logE),
             lambda: loss64(V64 - eps * dv, logE)),
            ("log E", float((g_logE * de).sum()), lambda: loss64(V64, logE + eps * de),
             lambda: loss64(V64, logE - eps * de))):
        fd = (plus() - minus()) / (2 * eps)
        print(f"{name}: finite difference {fd:.6e}, FP32 adjoint {an:.6e}, rel {abs(fd - an) / abs(fd):.2e}",
              flush=True)
        assert abs(fd - an) <= 1e-3 * abs(fd), (name, fd, an)
