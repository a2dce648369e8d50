"""Parity of the HIP step (through the C ABI) against the oracle.

Each case builds one seeded scene, runs it on the GPU and in the oracle from
identical inputs and compares particle X, V, F, B, valid and the grid mass and
velocity; the error of a quantity is max|a - b| / max|b|.

FP64 runs must match the FP64 oracle to 1e-9 after one step.  FP32 runs must
match, after one step, both the FP32 oracle and the FP64 oracle (the exact
answer for the same float inputs) to 1e-5 -- except B against the FP32
oracle: the reference evaluates B in a form that cancels O(|v||x|) terms to
produce an O(|v|h) result, so its FP32 B carries ~1e-4 relative rounding
error, while the kernel's B is checked against the FP64 oracle.  After ten
steps the FP32 drift from the FP64 oracle is bounded by 1e-4.
"""
import numpy as np
import pytest

import mpmtest as mt
from physbam_amd import mpmx

pytestmark = pytest.mark.gpu


def gpu_context(sc, prec=None):
    prec = prec or sc["prec"]
    d = sc["dim"]
    ctx = mpmx.Context(d, sc["N"], sc["lo"], sc["hi"], precision=prec)
    ctx.set_constitutive(sc["model"], sc.get("neo_threshold", 0.25))
    ctx.set_gravity(sc["gravity"])
    for w in sc["walls"]:
        ctx.add_collision_box(w["type"], w["friction"], w["lo"][:d], w["hi"][:d])
    ctx.set_particles(sc["X"], sc["V"], sc["F"], sc["B"], sc["mass"], sc["volume"], sc["mu"], sc["lambda"],
                      sc["valid"])
    return ctx


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def compare(gpu, orc, tol, tag, keys=("X", "V", "F", "B")):
    errs = {k: rel(gpu[k], orc[k]) for k in keys}
    print(tag, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert np.array_equal(gpu["valid"], orc["valid"]), tag
    for k, v in errs.items():
        assert v <= tol, (tag, k, v)
    return errs


sq2 = lambda X: (np.abs(X[:, 0] - 0.5) < 0.2) & (np.abs(X[:, 1] - 0.6) < 0.2)
sph3 = lambda X: ((X - np.array([0.5, 0.45, 0.5])) ** 2).sum(1) < 0.3 ** 2

CASES = [
    # dim, prec, model, wall type, friction
    (2, 8, mt.COROTATED, mt.STICK, 0.0),
    (2, 8, mt.NEO_HOOKEAN, mt.SLIP, 0.3),
    (2, 4, mt.COROTATED, mt.SEPARATE, 0.0),
    (3, 8, mt.COROTATED, mt.SLIP, 0.0),
    (3, 8, mt.NEO_HOOKEAN, mt.SEPARATE, 0.2),
    (3, 4, mt.COROTATED, mt.STICK, 0.0),
    (3, 4, mt.NEO_HOOKEAN, mt.SLIP, 0.0),
    (3, 4, mt.COROTATED, mt.SEPARATE, 0.3),
]


def scene(dim, prec, model, wtype, fric, N=None):
    N = N or (32 if dim == 2 else 16)
    walls = mt.wall_boxes(dim, [N] * dim, [0] * dim, [1] * dim, flags=(1 << (2 * dim)) - 1, inset=0.12,
                          type_=wtype, friction=fric)
    shape = sq2 if dim == 2 else sph3
    return mt.make_scene(dim, N, prec=prec, model=model, shape=shape, v0=[0.3, -3.0, 0.2][:dim], vrand=1.0,
                         brand=1e-3, frand=0.2, walls=walls, dt=1e-3, per_particle_E=True)


@pytest.mark.parametrize("dim,prec,model,wtype,fric", CASES)
def test_one_step_matches_oracle(dim, prec, model, wtype, fric):
    sc = scene(dim, prec, model, wtype, fric)
    tag = f"step1 d{dim} p{prec} m{model} w{wtype} f{fric} n={sc['X'].shape[0]}"
    print(tag, flush=True)
    orc = mt.oracle_run(sc, 1, prec=prec)
    with gpu_context(sc) as ctx:
        ctx.step(sc["dt"], 1)
        gpu = ctx.get_particles()
        gm, gv = ctx.get_grid()
    if prec == 8:
        compare(gpu, orc, 1e-9, tag)
    else:
        orc64 = mt.oracle_run(sc, 1, prec=8)
        compare(gpu, orc, 1e-5, tag + " vs fp32 oracle", keys=("X", "V", "F"))
        compare(gpu, orc64, 1e-5, tag + " vs fp64 oracle")
        print(tag, "fp32 oracle B vs fp64 oracle", f"{rel(orc['B'], orc64['B']):.2e}", flush=True)
    assert rel(gm, orc["grid_mass"]) <= (1e-12 if prec == 8 else 1e-5)
    assert rel(gv, orc["grid_velocity"]) <= (1e-9 if prec == 8 else 1e-5)


@pytest.mark.parametrize("dim,prec,model,wtype,fric", CASES)
def test_ten_steps_match_oracle(dim, prec, model, wtype, fric):
    sc = scene(dim, prec, model, wtype, fric)
    tag = f"step10 d{dim} p{prec} m{model} w{wtype} f{fric}"
    print(tag, flush=True)
    orc64 = mt.oracle_run(sc, 10, prec=8, want_grid=False)
    with gpu_context(sc) as ctx:
        ctx.set_reorder_interval(3)
        ctx.step(sc["dt"], 10)
        gpu = ctx.get_particles()
    if prec == 8:
        compare(gpu, orc64, 1e-8, tag)
    else:
        orc = mt.oracle_run(sc, 10, prec=4, want_grid=False)
        e = compare(gpu, orc64, 1e-4, tag + " vs fp64 oracle")
        r = {k: rel(orc[k], orc64[k]) for k in e}
        print(tag, "fp32 oracle vs fp64 oracle", {k: f"{v:.2e}" for k, v in r.items()}, flush=True)


@pytest.mark.parametrize("threads,wg", [(256, 0), (128, 512), (256, 512)])
@pytest.mark.parametrize("dim,prec,model,wtype,fric", [CASES[1], CASES[4], CASES[7]])
def test_ten_steps_with_other_launch_configs(dim, prec, model, wtype, fric, threads, wg):
    """Ten fused steps with the fused kernel at 256 threads per workgroup
    and/or a fixed count of workgroups striding over the tile list, against
    the FP64 oracle to the bounds of the default configuration (2D and 3D,
    FP64 and FP32, friction)."""
    sc = scene(dim, prec, model, wtype, fric)
    tag = f"launch t{threads} w{wg} d{dim} p{prec} m{model} w{wtype} f{fric}"
    print(tag, flush=True)
    orc64 = mt.oracle_run(sc, 10, prec=8, want_grid=False)
    with gpu_context(sc) as ctx:
        ctx.set_launch_config(graph=False, tile_workgroups=wg, tile_threads=threads)
        ctx.set_reorder_interval(3)
        ctx.step(sc["dt"], 10)
        gpu = ctx.get_particles()
    compare(gpu, orc64, 1e-8 if prec == 8 else 1e-4, tag)


# Bins are rebuilt only when enough particles have left their binned tile;
# until then those particles go through the mover kernels.  Each policy must
# reproduce the oracle: never rebuilding (every tile change is a mover),
# rebuilding every step (no movers), and the default threshold.
POLICIES = [("never", 2.0, 0), ("every_step", 1.0 / 128, 1), ("default", 1.0 / 128, 0)]


def mover_scene(dim, prec, walls=True):
    N = 32 if dim == 2 else 16
    w = (mt.wall_boxes(dim, [N] * dim, [0] * dim, [1] * dim, flags=(1 << (2 * dim)) - 1, inset=0.12,
                       type_=mt.SLIP, friction=0.0) if walls else [])
    shape = sq2 if dim == 2 else sph3
    return mt.make_scene(dim, N, prec=prec, model=mt.COROTATED, shape=shape, v0=[2.0, -5.0, 1.5][:dim],
                         vrand=4.0, brand=1e-3, frand=0.1, walls=w, dt=1e-3, per_particle_E=True)


@pytest.mark.parametrize("policy,frac,maxint", POLICIES)
@pytest.mark.parametrize("dim,prec", [(2, 8), (3, 8), (3, 4)])
def test_movers_between_rebins(dim, prec, policy, frac, maxint):
    sc = mover_scene(dim, prec)
    steps = 20
    tag = f"movers {policy} d{dim} p{prec} n={sc['X'].shape[0]}"
    print(tag, flush=True)
    orc64 = mt.oracle_run(sc, steps, prec=8, want_grid=False)
    with gpu_context(sc) as ctx:
        ctx.set_rebin_policy(frac, maxint)
        ctx.set_reorder_interval(4)
        movers = 0
        for _ in range(steps):
            ctx.step(sc["dt"], 1)
            movers += ctx.bin_stats()[1]
        rebins = ctx.bin_stats()[0]
        gpu = ctx.get_particles()
    print(tag, "rebins", rebins, "mover-steps", movers, flush=True)
    if policy == "never":
        assert rebins == 1 and movers > 0
    if policy == "every_step":
        assert rebins == steps and movers == 0
    compare(gpu, orc64, 1e-8 if prec == 8 else 1e-4, tag)


@pytest.mark.parametrize("dim,prec", [(2, 8), (3, 4)])
def test_particles_leaving_the_domain_between_rebins(dim, prec):
    # no walls: the body falls through y = 0 and its particles are invalidated
    # (MPM_DRIVER.cpp:1227-1238) while the bins are never rebuilt
    sc = mover_scene(dim, prec, walls=False)
    sc["X"][:, 1] -= 0.3 if dim == 2 else 0.25
    steps = 25
    orc64 = mt.oracle_run(sc, steps, prec=8, want_grid=False)
    with gpu_context(sc) as ctx:
        ctx.set_rebin_policy(2.0, 0)
        ctx.step(sc["dt"], steps)
        gpu = ctx.get_particles()
    lost = int((gpu["valid"] == 0).sum())
    print("leaving", dim, prec, "invalidated", lost, flush=True)
    assert lost > 0
    compare(gpu, orc64, 1e-8 if prec == 8 else 1e-4, f"leaving d{dim} p{prec}")


def test_sort_is_stable_and_exact():
    rng = np.random.default_rng(7)
    with mpmx.Context(3, 8, [0, 0, 0], [1, 1, 1]) as ctx:
        for n, bits in ((1, 4), (1000, 5), (100003, 19), (262144, 8)):
            keys = rng.integers(0, (1 << bits) - 1, n).astype(np.uint32)
            keys[rng.random(n) < 0.05] = 0xFFFFFFFF
            print("sort", n, bits, flush=True)
            ko, order = ctx.sort_keys(keys, bits)
            ref = np.argsort(keys, kind="stable")
            assert np.array_equal(order, ref)
            assert np.array_equal(ko, keys[ref])


def test_svd_kernel():
    rng = np.random.default_rng(3)
    n = 4096
    F = np.tile(np.eye(3).reshape(-1), (n, 1)) + 0.4 * (rng.random((n, 9)) - 0.5)
    F[:8] = [[1, 0, 0, 0, 1, 0, 0, 0, 1], [2, 0, 0, 0, 2, 0, 0, 0, 2], [1, 0, 0, 0, 1, 0, 0, 0, -1],
             [0, 1, 0, 1, 0, 0, 0, 0, 1], [1e-4, 0, 0, 0, 1, 0, 0, 0, 1], [3, 0, 0, 0, 1, 0, 0, 0, 1e-6],
             [1, 1, 0, 0, 1, 0, 0, 0, 1], [-1, 0, 0, 0, -1, 0, 0, 0, -1]]
    with mpmx.Context(3, 8, [0, 0, 0], [1, 1, 1], precision=mpmx.FP64) as ctx:
        U, S, V = ctx.svd(F)
    Um = U.reshape(n, 3, 3).transpose(0, 2, 1)
    Vm = V.reshape(n, 3, 3).transpose(0, 2, 1)
    Fm = F.reshape(n, 3, 3).transpose(0, 2, 1)
    rec = Um @ (S[:, :, None] * Vm.transpose(0, 2, 1))
    assert np.abs(rec - Fm).max() < 1e-12
    assert np.abs(Um @ Um.transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(Um) - 1).max() < 1e-12 and np.abs(np.linalg.det(Vm) - 1).max() < 1e-12
    a = np.abs(S)
    assert (a[:, 0] >= a[:, 1] - 1e-14).all() and (a[:, 1] >= a[:, 2] - 1e-14).all()
    assert (S[:, :2] >= 0).all()
    _, So, _, _ = mt.oracle_svd(F, 3)
    assert np.abs(S - So).max() < 1e-12


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("dim", [2, 3])
def test_svd_factors_match_the_oracle(dim, prec):
    """mpmx_svd in 2D and 3D, FP32 and FP64: the reference's convention
    (MATRIX_3X3.cpp:198-236, MATRIX_2X2.h:313-322: U, V rotations, |s| sorted
    descending, only the last singular value negative), reconstruction and
    orthonormality to the precision, and U, s, V equal to the oracle's (the
    reference's own algorithm, same precision input) where the singular values
    are separated by more than 5 % (then the factors are unique up to the
    convention's freedom of negating the same columns of U and V in pairs, or
    both whole factors in 2D; closer singular values leave U and V free inside
    the near-degenerate subspace)."""
    rng = np.random.default_rng(17 + dim)
    n = 4096
    d = dim
    F = np.tile(np.eye(d).reshape(-1), (n, 1)) + 0.8 * (rng.random((n, d * d)) - 0.5)
    F[: n // 8] *= -1  # inverted ones (det < 0 in 3D)
    eps = 1e-12 if prec == 8 else 3e-6
    with mpmx.Context(d, 8, [0] * d, [1] * d, precision=prec) as ctx:
        U, S, V = ctx.svd(F)
    dt_np = np.float64 if prec == 8 else np.float32
    Fin = F.astype(dt_np).astype(np.float64)
    Um = U.astype(np.float64).reshape(n, d, d).transpose(0, 2, 1)
    Vm = V.astype(np.float64).reshape(n, d, d).transpose(0, 2, 1)
    Fm = Fin.reshape(n, d, d).transpose(0, 2, 1)
    S = S.astype(np.float64)
    rec = Um @ (S[:, :, None] * Vm.transpose(0, 2, 1))
    assert np.abs(rec - Fm).max() <= eps * 4 * np.abs(Fm).max()
    for Q in (Um, Vm):
        assert np.abs(Q @ Q.transpose(0, 2, 1) - np.eye(d)).max() <= 4 * eps
        assert np.abs(np.linalg.det(Q) - 1).max() <= 4 * eps
    a = np.abs(S)
    for i in range(d - 1):
        assert (a[:, i] >= a[:, i + 1] - 4 * eps * a[:, 0]).all()
        assert (S[:, i] >= 0).all()
    Uo, So, Vo, _ = mt.oracle_svd(F, d, prec=prec)
    So = So.astype(np.float64)
    assert np.abs(S - So).max() <= 8 * eps * np.abs(So).max()
    gaps = np.min(np.abs(np.diff(np.abs(So), axis=1)), axis=1) / np.abs(So[:, 0])
    sep = (gaps > 0.05) & (np.abs(So[:, -1]) > 0.05 * np.abs(So[:, 0]))
    assert sep.sum() > n // 4
    # U diag(s) and V diag(s) with s = +-1 per column and prod s = 1 factor
    # F the same way under the convention: align the columns' signs on U,
    # then V must match with the same signs
    Ug, Ur = Um[sep], Uo.astype(np.float64).reshape(n, d, d).transpose(0, 2, 1)[sep]
    Vg, Vr = Vm[sep], Vo.astype(np.float64).reshape(n, d, d).transpose(0, 2, 1)[sep]
    sg = np.sign(np.einsum("nij,nij->nj", Ug, Ur))
    assert np.all(np.prod(sg, axis=1) == 1)
    for got, ref in ((Ug * sg[:, None, :], Ur), (Vg * sg[:, None, :], Vr)):
        err = np.abs(got - ref).max()
        print(f"dim {d} prec {prec}: max factor difference {err:.2e} over {int(sep.sum())} matrices "
              f"({int((sg < 0).any(axis=1).sum())} with flipped column pairs)", flush=True)
        assert err <= (1e-10 if prec == 8 else 2e-4)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("dim", [2, 3])
def test_stress_near_rotations(dim, prec):
    # F = Q (I + e S): strains on both sides of the small-strain series radius
    rng = np.random.default_rng(5)
    per = 256
    Fs = []
    for e in (1e-7, 1e-4, 1e-2, 0.05, 0.1, 0.12, 0.13, 0.2, 0.4):
        A = rng.standard_normal((per, dim, dim))
        Q, _ = np.linalg.qr(A)
        Q[:, :, -1] *= np.sign(np.linalg.det(Q))[:, None]  # proper rotations (det +1) in 2D and 3D
        S = rng.standard_normal((per, dim, dim))
        S = 0.5 * (S + S.transpose(0, 2, 1))
        S /= np.linalg.norm(S, axis=(1, 2), keepdims=True)
        M = Q @ (np.eye(dim) + e * S)
        Fs.append(M.transpose(0, 2, 1).reshape(per, -1))  # column-major
    F = np.concatenate(Fs)
    n = F.shape[0]
    vol = np.full(n, 1.5e-6)
    mu, la = mt.lame(rng.uniform(1e3, 1e5, n), rng.uniform(0.2, 0.4, n))
    dt = np.float32 if prec == 4 else np.float64
    F, vol, mu, la = (np.asarray(x, dt).astype(np.float64) for x in (F, vol, mu, la))
    with mpmx.Context(dim, 8, [0] * dim, [1] * dim, precision=prec) as ctx:
        ctx.set_constitutive(mt.COROTATED, 0.25)
        tau = ctx.stress(F, vol, mu, la)
    ref = mt.oracle_stress(F, vol, mu, la, dim, mt.COROTATED, 0.25, prec=8)
    # Near a rotation F - I is O(1) while the strain is tiny: rounding in F F^T
    # is O(eps) in units of the modulus, so the scale has a floor of
    # 5% of vol (2 mu + lambda) (errors below ~1e-6 of the modulus pass).
    scale = np.maximum(np.abs(ref).max(axis=1), 0.05 * vol * (2 * mu + la))[:, None]
    err = (np.abs(tau - ref) / scale).max(axis=1)
    for i, e in enumerate((1e-7, 1e-4, 1e-2, 0.05, 0.1, 0.12, 0.13, 0.2, 0.4)):
        print("stress near rotation", dim, prec, e, f"{err[i * per:(i + 1) * per].max():.2e}", flush=True)
    assert err.max() < (1e-10 if prec == 8 else 2e-5)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("model", [mt.COROTATED, mt.HENCKY])
@pytest.mark.parametrize("dim", [2, 3])
def test_stress_repeated_stretches(dim, model, prec):
    """Stress past the series radius where the stretches of F repeat: F = s R
    (all equal: the characteristic polynomial of F F^T - I has a root of
    multiplicity d), in 3D F = R diag(a, a, b) (a double root), and nearly
    repeated stretches, against the FP64 oracle; FP32 to 2e-5 of the largest
    entry, as the other stress tests."""
    rng = np.random.default_rng(17)
    if dim == 3:
        sets = ([1.3, 1.3, 1.3], [0.8, 0.8, 0.8], [1.25, 1.25, 0.9], [0.85, 1.2, 1.2],
                [1.2, 1.2 * (1 + 1e-6), 0.9], [1.5, 0.7, 0.7 * (1 - 1e-5)])
    else:
        sets = ([1.3, 1.3], [0.8, 0.8], [1.2, 1.2 * (1 + 1e-6)], [0.7, 0.7 * (1 - 1e-5)])
    mats = []
    for _ in range(64):
        R = np.linalg.qr(rng.standard_normal((dim, dim)))[0]
        R[:, -1] *= np.sign(np.linalg.det(R))
        Q = np.linalg.qr(rng.standard_normal((dim, dim)))[0]
        for st in sets:
            mats.append(R @ Q @ np.diag(st) @ Q.T)
    F = np.stack(mats).transpose(0, 2, 1).reshape(len(mats), dim * dim)  # column-major
    n = F.shape[0]
    vol = np.full(n, 1.5e-6)
    mu, la = mt.lame(rng.uniform(1e3, 1e5, n), rng.uniform(0.2, 0.4, n))
    dt = np.float32 if prec == 4 else np.float64
    F, vol, mu, la = (np.asarray(x, dt).astype(np.float64) for x in (F, vol, mu, la))
    with mpmx.Context(dim, 8, [0] * dim, [1] * dim, precision=prec) as ctx:
        ctx.set_constitutive(model, 0.25)
        tau = ctx.stress(F, vol, mu, la)
    ref = mt.oracle_stress(F, vol, mu, la, dim, model, 0.25, prec=8)
    err = (np.abs(tau - ref) / np.abs(ref).max(axis=1, keepdims=True)).max()
    print("repeated stretches", dim, model, prec, f"{err:.2e}", flush=True)
    assert err < (1e-10 if prec == 8 else 2e-5)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("dim", [2, 3])
def test_stress_kernel(dim, model, prec):
    rng = np.random.default_rng(11)
    n = 2048
    F = np.tile(np.eye(dim).reshape(-1), (n, 1)) + 0.3 * (rng.random((n, dim * dim)) - 0.5)
    # compressed particles reach the neo-Hookean clamp branch (J < 0.25)
    F[: n // 8] *= 0.5
    vol = rng.uniform(1e-6, 2e-6, n)
    mu, la = mt.lame(rng.uniform(1e3, 1e5, n), rng.uniform(0.2, 0.4, n))
    dt = np.float32 if prec == 4 else np.float64
    F, vol, mu, la = (np.asarray(x, dt).astype(np.float64) for x in (F, vol, mu, la))
    with mpmx.Context(dim, 8, [0] * dim, [1] * dim, precision=prec) as ctx:
        ctx.set_constitutive(model, 0.25)
        tau = ctx.stress(F, vol, mu, la)
    ref = mt.oracle_stress(F, vol, mu, la, dim, model, 0.25, prec=8)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = (np.abs(tau - ref) / scale).max()
    print("stress", dim, model, prec, err, flush=True)
    assert err < (1e-10 if prec == 8 else 2e-5)


@pytest.mark.parametrize("prec", [4, 8])
def test_no_particles_and_all_invalid(prec):
    """A context without particles steps (time advances, nothing else), and a
    scene whose particles are all invalid steps through the fused and unfused
    paths with an empty tile list, leaving every particle where it was."""
    sc = scene(3, prec, mt.COROTATED, mt.STICK, 0.0)
    d = 3
    with mpmx.Context(d, sc["N"], sc["lo"], sc["hi"], precision=prec) as ctx:
        ctx.set_gravity(sc["gravity"])
        z = lambda *s: np.zeros(s)
        ctx.set_particles(z(0, 3), z(0, 3), z(0, 9), z(0, 9), z(0), z(0), z(0), z(0), np.zeros(0, np.uint8))
        ctx.step(sc["dt"], 3)
        assert ctx.time == pytest.approx(3 * sc["dt"])
        assert ctx.get_particles()["X"].shape[0] == 0
    sc0 = dict(sc, valid=np.zeros_like(sc["valid"]))
    for fused in (True, False):
        with gpu_context(sc0) as ctx:
            if not fused:
                ctx.enable_grid_stats()  # grid statistics run the unfused step
            ctx.step(sc["dt"], 4)
            got = ctx.get_particles()
            assert not got["valid"].any()
            for k in ("X", "V", "F", "B"):
                assert np.array_equal(got[k], np.asarray(sc0[k], got[k].dtype)), (fused, k)
            gm, _ = ctx.get_grid()
            assert not np.any(gm), fused


@pytest.mark.parametrize("dim,prec", [(2, 8), (3, 8), (3, 4)])
def test_single_particle_matches_oracle(dim, prec):
    """One particle (a tile holding a single particle, no other tile active),
    five steps against the oracle."""
    sc = scene(dim, prec, mt.COROTATED, mt.SLIP, 0.2)
    i = int(np.argmax(sc["X"][:, 1]))  # the highest particle: it falls freely for a while
    sc1 = dict(sc)
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc1[k] = np.ascontiguousarray(np.asarray(sc[k])[i:i + 1])
    orc = mt.oracle_run(sc1, 5, prec=8, want_grid=False)
    with gpu_context(sc1) as ctx:
        ctx.step(sc["dt"], 5)
        gpu = ctx.get_particles()
    compare(gpu, orc, 1e-9 if prec == 8 else 1e-5, f"single particle d{dim} p{prec}")


@pytest.mark.parametrize("dim,prec,threads", [(3, 4, 128), (3, 4, 256), (3, 8, 128), (3, 8, 256), (2, 4, 128),
                                             (2, 8, 256)])
def test_crowded_cell_matches_oracle(dim, prec, threads):
    """1500 particles inside one grid cell plus a sparse cloud around it: the
    fused kernel's per-cell ranks run far past the wavefronts of a workgroup
    (hundreds of rounds), three steps against the FP64 oracle."""
    sc = scene(dim, prec, mt.COROTATED, mt.STICK, 0.0)
    rng = np.random.default_rng(23)
    h = 1.0 / sc["N"][0]
    m = 1500
    Xc = np.full(dim, 0.5) + (0.1 + 0.8 * rng.random((m, dim))) * h  # one cell
    keep = np.arange(0, sc["X"].shape[0], 7)                          # a sparse cloud
    cat = lambda a, b: np.concatenate([np.asarray(a), np.asarray(b)]).astype(np.asarray(a).dtype)
    sc1 = dict(sc)
    sc1["X"] = cat(sc["X"][keep], Xc)
    sc1["V"] = cat(sc["V"][keep], 0.2 * rng.standard_normal((m, dim)))
    sc1["F"] = cat(sc["F"][keep], np.tile(np.eye(dim).reshape(-1), (m, 1)) + 0.01 * rng.standard_normal((m, dim * dim)))
    sc1["B"] = cat(sc["B"][keep], np.zeros((m, dim * dim)))
    for k in ("mass", "volume", "mu", "lambda"):
        sc1[k] = cat(sc[k][keep], np.full(m, np.asarray(sc[k]).mean()))
    sc1["valid"] = cat(sc["valid"][keep], np.ones(m, np.uint8))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda"):
        sc1[k] = np.asarray(sc1[k], np.float32 if prec == 4 else np.float64)
    orc = mt.oracle_run(sc1, 3, prec=8, want_grid=False)
    with gpu_context(sc1) as ctx:
        ctx.set_launch_config(tile_threads=threads)
        ctx.step(sc["dt"], 3)
        gpu = ctx.get_particles()
    compare(gpu, orc, 1e-9 if prec == 8 else 1e-5, f"crowded cell d{dim} p{prec} t{threads}")


@pytest.mark.parametrize("dim,prec", [(3, 4), (3, 8), (2, 4), (2, 8)])
def test_fast_particles_cross_tiles(dim, prec):
    """A body translating by (1.5, -2, 1) cells per step and spinning at up
    to two cells per step at its rim (no CFL limit, as the reference allows
    with a fixed dt): with the bins never rebuilt, the particles cross whole
    tiles (about five cells in three steps), most become movers and a few
    leave the domain; three fused steps against the FP64 oracle."""
    sc = scene(dim, prec, mt.COROTATED, mt.SLIP, 0.0)
    h, dt = 1.0 / sc["N"][0], sc["dt"]
    X = np.asarray(sc["X"], np.float64)
    r = X - X.mean(axis=0)
    rmax = np.linalg.norm(r, axis=1).max()
    if dim == 3:
        w = np.array([0.3, 1.0, -0.5])
        w *= 2.0 * h / dt / (np.linalg.norm(w) * rmax)
        spin = np.cross(w, r)
    else:
        spin = (2.0 * h / dt / rmax) * np.stack([-r[:, 1], r[:, 0]], axis=1)
    V = np.array([1.5, -2.0, 1.0][:dim]) * h / dt + spin
    sc1 = dict(sc, V=np.asarray(V, np.asarray(sc["V"]).dtype))
    orc = mt.oracle_run(sc1, 3, prec=8, want_grid=False)
    print("fast particles: valid after 3 steps", int(orc["valid"].sum()), "of", X.shape[0], flush=True)
    with gpu_context(sc1) as ctx:
        ctx.set_rebin_policy(2.0, 0)  # never rebuild: every tile change stays a mover
        ctx.step(dt, 3)
        gpu = ctx.get_particles()
    compare(gpu, orc, 1e-9 if prec == 8 else 1e-5, f"fast particles d{dim} p{prec}")
