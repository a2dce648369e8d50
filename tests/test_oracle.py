"""The oracle pinned against the reference and against the mathematics.

CPU only.  Three kinds of checks:
  * golden vectors written by the PhysBAM reference itself (tests/golden/,
    made by make_golden.py): the oracle must reproduce them bit for bit,
    since it restates the reference in the same operation order;
  * the live reference harness, where it is built (authoring container);
  * closed forms and invariants that a plausible mistake would break:
    partition of unity and linear reproduction of the weights, SVD
    reconstruction / rotations / ordering, P_hat = dPsi/dsigma by finite
    differences of the textbook energies, mass and momentum conservation of
    P2G, exact transfer of an affine velocity field.
"""
import json
import os

import numpy as np
import pytest

import mpmtest as mt
import scene_io

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def scene_from_fixture(z):
    meta = json.loads(str(z["meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z["walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z["in_" + k]
    return sc


STEP_FIXTURES = ["step_square2d_f64.npz", "step_sphere3d_neo_slip_f32.npz", "step_sphere3d_coro_sep_fric_f64.npz",
                 "step_square2d_neo_clamp_f64.npz"]


@pytest.mark.parametrize("fixture", STEP_FIXTURES)
def test_oracle_reproduces_reference_steps_bit_exact(fixture):
    z = load(fixture)
    sc = scene_from_fixture(z)
    steps = sorted({int(k[1:].split("_")[0]) for k in z.files if k.startswith("s") and k[1].isdigit()})
    for n in steps:
        orc = mt.oracle_run(sc, n, want_grid=True)
        for k in ("X", "V", "F", "B", "valid"):
            assert np.array_equal(orc[k].astype(np.float64) if k != "valid" else orc[k], z[f"s{n}_{k}"]), (fixture, n, k)
        if f"s{n}_grid_mass" in z.files:
            assert np.array_equal(orc["grid_mass"].astype(np.float64), z[f"s{n}_grid_mass"])
            for k in ("grid_velocity", "grid_velocity_save", "grid_velocity_friction_save"):
                assert np.array_equal(orc[k].astype(np.float64), z[f"s{n}_{k}"]), (fixture, n, k)


def test_clamp_fixture_reaches_the_clamp_branch():
    z = load("step_square2d_neo_clamp_f64.npz")
    F = z["in_F"].reshape(-1, 2, 2)
    assert (np.linalg.det(F) < 0.25).sum() > 50


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", [8, 4])
def test_oracle_svd_matches_reference(dim, prec):
    z = load("components.npz")
    A = z[f"svd{dim}_p{prec}_A"]
    U, S, V, _ = mt.oracle_svd(A, dim, prec=prec)
    for got, key in ((U, "U"), (S, "S"), (V, "V")):
        assert np.array_equal(got.astype(np.float64), z[f"svd{dim}_p{prec}_{key}"]), key


@pytest.mark.parametrize("dim", [2, 3])
def test_oracle_svd_properties(dim):
    rng = np.random.default_rng(1)
    n = 2000
    A = rng.normal(size=(n, dim * dim)) * rng.choice([1e-4, 1.0, 50.0], size=(n, 1))
    U, S, V, it = mt.oracle_svd(A, dim)
    Um = U.reshape(n, dim, dim).transpose(0, 2, 1)
    Vm = V.reshape(n, dim, dim).transpose(0, 2, 1)
    Am = A.reshape(n, dim, dim).transpose(0, 2, 1)
    rec = Um @ (S[:, :, None] * Vm.transpose(0, 2, 1))
    # The reference deflates with tau = 16 eps max(|B|_F / 2, 1) (Core/Matrices/MATRIX_3X3.cpp:213-214),
    # an absolute floor for small matrices, so errors scale with max(|A|, 1).
    scale = np.maximum(np.abs(Am).max(axis=(1, 2)), 1.0)
    assert (np.abs(rec - Am).max(axis=(1, 2)) <= 1e-12 * scale).all()
    assert np.abs(Um @ Um.transpose(0, 2, 1) - np.eye(dim)).max() < 1e-12
    assert np.abs(np.linalg.det(Um) - 1).max() < 1e-12
    assert np.abs(np.linalg.det(Vm) - 1).max() < 1e-12
    a = np.abs(S)
    for i in range(dim - 1):
        assert (a[:, i] >= a[:, i + 1]).all()
        assert (S[:, i] >= 0).all()
    assert (np.abs(a - np.linalg.svd(Am, compute_uv=False)) <= 1e-12 * scale[:, None]).all()
    assert (np.sign(S[:, -1]) == np.sign(np.linalg.det(Am))).all()
    if dim == 3:
        assert it.max() < 100


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("prec", [8, 4])
def test_oracle_p_hat_matches_reference(dim, model, prec):
    z = load("components.npz")
    k = f"pstrain{dim}_m{model}_p{prec}"
    P = mt.oracle_pstrain(z[k + "_S"], z[k + "_mu"], z[k + "_la"], dim, model, 0.25, prec=prec)
    assert np.array_equal(P.astype(np.float64), z[k + "_P"])


def energy(S, mu, la, model, thr=0.25):
    """Textbook energy densities: fixed corotated mu sum (s-1)^2 + la/2 (J-1)^2;
    neo-Hookean mu/2 (sum s^2 - d) - mu ln J + la/2 (ln J)^2 (J >= threshold);
    St. Venant-Kirchhoff with Hencky strain mu sum (ln s)^2 + la/2 (sum ln s)^2."""
    J = np.prod(S, axis=1)
    if model == 0:
        return mu * ((S - 1) ** 2).sum(1) + la / 2 * (J - 1) ** 2
    if model == 2:
        L = np.log(S)
        return mu * (L ** 2).sum(1) + la / 2 * L.sum(1) ** 2
    d = S.shape[1]
    lj = np.log(J)
    return mu / 2 * ((S ** 2).sum(1) - d) - mu * lj + la / 2 * lj ** 2


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("model", [0, 1, 2])
def test_p_hat_is_energy_gradient(dim, model):
    rng = np.random.default_rng(5)
    n = 300
    S = rng.uniform(0.7, 1.3, (n, dim))
    mu, la = mt.lame(rng.uniform(1e3, 1e5, n), rng.uniform(0.2, 0.4, n))
    P = mt.oracle_pstrain(S, mu, la, dim, model)
    h = 1e-6
    for i in range(dim):
        Sp, Sm = S.copy(), S.copy()
        Sp[:, i] += h
        Sm[:, i] -= h
        fd = (energy(Sp, mu, la, model) - energy(Sm, mu, la, model)) / (2 * h)
        assert np.allclose(P[:, i], fd, rtol=1e-6, atol=1e-6 * (mu + la))


def test_neo_hookean_clamp_is_continuous_at_threshold():
    mu, la = mt.lame(1e4, 0.3)
    # J = 0.25 exactly along the diagonal: both branches must agree there
    s = 0.25 ** (1 / 3)
    eps = 1e-9
    S = np.array([[s + eps] * 3, [s - eps] * 3])
    P = mt.oracle_pstrain(S, mu, la, 3, 1)
    assert np.allclose(P[0], P[1], rtol=1e-6)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", [8, 4])
def test_oracle_weights_match_reference(dim, prec):
    z = load("components.npz")
    X = z[f"weights{dim}_p{prec}_X"]
    base, w, dw = mt.oracle_weights(X, dim, 16, 0.0, 1.0, prec=prec)
    assert np.array_equal(base, z[f"weights{dim}_p{prec}_base"])
    assert np.array_equal(w.astype(np.float64), z[f"weights{dim}_p{prec}_w"])
    assert np.array_equal(dw.astype(np.float64), z[f"weights{dim}_p{prec}_dw"])


@pytest.mark.parametrize("dim", [2, 3])
def test_weights_partition_and_linear_reproduction(dim):
    rng = np.random.default_rng(9)
    N = 16
    h = 1.0 / N
    X = rng.uniform(0.1, 0.9, (500, dim))
    base, w, dw = mt.oracle_weights(X, dim, N, 0.0, 1.0)
    grids = np.stack(np.meshgrid(*[np.arange(3)] * dim, indexing="ij"), -1).reshape(-1, dim)
    for p in range(len(X)):
        W = np.prod([w[p, grids[:, a], a] for a in range(dim)], axis=0)
        G = np.stack([dw[p, grids[:, j], j] * np.prod([w[p, grids[:, a], a] for a in range(dim) if a != j], axis=0)
                      for j in range(dim)], axis=1)
        xi = (base[p] + grids + 0.5) * h
        d = xi - X[p]
        assert abs(W.sum() - 1) < 1e-14
        assert np.abs((W[:, None] * d).sum(0)).max() < 1e-15
        assert np.abs(G.sum(0)).max() < 1e-12
        assert np.abs(d.T @ G - np.eye(dim)).max() < 1e-12            # sum (x_i - x_p) grad w^T = I
        assert np.allclose((W[:, None, None] * d[:, :, None] * d[:, None, :]).sum(0), np.eye(dim) * h * h / 4)


def test_p2g_conserves_mass_and_momentum():
    sq = lambda X: (np.abs(X[:, 0] - 0.5) < 0.2) & (np.abs(X[:, 1] - 0.5) < 0.2)
    sc = mt.make_scene(2, 32, prec=8, shape=sq, vrand=2.0, brand=1e-3, gravity=[0.0, 0.0], dt=1e-4)
    out = mt.oracle_run(sc, 1)
    # grid mass after P2G equals particle mass; P2G momentum (velocity_save) equals particle momentum
    assert np.isclose(out["grid_mass"].sum(), sc["mass"].sum(), rtol=1e-13)
    gp = (out["grid_mass"][:, None] * out["grid_velocity_save"]).sum(0)
    pp = (sc["mass"][:, None] * sc["V"]).sum(0)
    assert np.allclose(gp, pp, rtol=1e-12, atol=1e-14)
    # without walls or gravity internal forces cancel: the updated grid momentum is unchanged
    gv = (out["grid_mass"][:, None] * out["grid_velocity"]).sum(0)
    assert np.allclose(gv, pp, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_affine_velocity_field_is_transferred_exactly(dim):
    """v(x) = v0 + A x carried by B = A D_p (D_p = h^2/4 I): P2G+G2P returns
    v_p and grad v = A; with dt -> tiny and F = I, F becomes I + dt A."""
    rng = np.random.default_rng(3)
    N = 16
    h = 1.0 / N
    A = rng.normal(size=(dim, dim))
    v0 = rng.normal(size=dim)
    box = lambda X: (np.abs(X - 0.5) < 0.25).all(1)
    sc = mt.make_scene(dim, N, prec=8, shape=box, gravity=[0.0] * dim, dt=1e-9, E=1e-30, nu=0.3)
    X = sc["X"]
    sc["V"] = v0 + X @ A.T
    Bm = A * h * h / 4
    sc["B"] = np.tile(Bm.T.reshape(-1), (len(X), 1))  # column-major
    out = mt.oracle_run(sc, 1)
    Xn = X + sc["dt"] * sc["V"]
    assert np.allclose(out["V"], v0 + X @ A.T, rtol=0, atol=1e-10)
    assert np.allclose(out["X"], Xn, rtol=0, atol=1e-14)
    Fexp = (np.eye(dim) + sc["dt"] * A).T.reshape(-1)
    assert np.allclose(out["F"], Fexp, rtol=0, atol=1e-14)
    assert np.allclose(out["B"], sc["B"], rtol=1e-8, atol=1e-14)


def test_free_fall_is_exact():
    sq = lambda X: (np.abs(X[:, 0] - 0.5) < 0.1) & (np.abs(X[:, 1] - 0.6) < 0.1)
    sc = mt.make_scene(2, 32, prec=8, shape=sq, dt=1e-3)
    out = mt.oracle_run(sc, 5)
    assert np.allclose(out["V"][:, 1], -9.8 * 5e-3, rtol=1e-12)
    assert np.allclose(out["V"][:, 0], 0, atol=1e-14)


def test_particles_leaving_the_domain_are_invalidated():
    sq = lambda X: (np.abs(X[:, 0] - 0.5) < 0.1) & (np.abs(X[:, 1] - 0.5) < 0.1)
    sc = mt.make_scene(2, 16, prec=8, shape=sq, gravity=[0.0, 0.0], dt=1e-3)
    sc["X"][0] = [0.99999, 0.5]
    sc["V"][0] = [100.0, 0.0]
    sc["valid"][1] = 0
    out = mt.oracle_run(sc, 1)
    assert out["valid"][0] == 0 and out["valid"][1] == 0 and out["valid"][2:].all()
    assert np.array_equal(out["X"][1], sc["X"][1])  # invalid particles are not simulated


@pytest.mark.skipif(not scene_io.ref_available(), reason="reference harness not built here")
def test_oracle_matches_live_reference_harness():
    sph = lambda X: ((X - np.array([0.5, 0.5, 0.5])) ** 2).sum(1) < 0.3 ** 2
    walls = mt.wall_boxes(3, [12] * 3, [0] * 3, [1] * 3, flags=63, inset=0.15, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(3, 12, prec=4, model=mt.NEO_HOOKEAN, shape=sph, v0=[1.0, -2.0, 0.5], vrand=0.5,
                       brand=1e-3, frand=0.3, walls=walls, dt=5e-4, seed=99)
    ref = scene_io.run_ref(sc, 8)[-1]
    orc = mt.oracle_run(sc, 8)
    for k in ("X", "V", "F", "B"):
        assert np.array_equal(orc[k].astype(np.float64), ref[k]), k


def frame_scene(z, name):
    meta = json.loads(str(z[f"{name}_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[f"{name}_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    return sc, meta


@pytest.mark.parametrize("name", ["throw2d_f64", "ball3d_neo_slip_f32"])
def test_oracle_adaptive_frames_match_reference(name):
    """Compute_Dt + clamp + the frame-end rule of Simulate_To_Frame: the
    substep dts and the final state agree bit for bit with the reference."""
    z = load("frames.npz")
    sc, m = frame_scene(z, name)
    out, dts = mt.oracle_frames(sc, m["nframes"], m["frame_dt"], m["cfl"], m["min_dt"], m["max_dt"])
    ref = z[f"{name}_dts"]
    assert len(dts) == len(ref)
    assert np.array_equal(dts, ref)
    for k in ("X", "V", "F", "B", "valid"):
        assert np.array_equal(out[k].astype(np.float64) if k != "valid" else out[k], z[f"{name}_out_{k}"]), k
    # the fixture exercises the CFL branch, the max_dt clamp and the frame-end split
    assert dts.min() < 0.999 * m["max_dt"]
    assert np.isclose(dts.sum(), m["nframes"] * m["frame_dt"], rtol=1e-5)


def test_grid_v_upper_bound_closed_form():
    rng = np.random.default_rng(3)
    for dim in (2, 3):
        n = 1000
        sc = dict(dim=dim, N=[40, 20, 10][:dim], lo=[0.0] * dim, hi=[1.0, 0.5, 0.5][:dim],
                  V=rng.normal(size=(n, dim)), B=rng.normal(size=(n, dim * dim)) * 1e-3,
                  valid=(rng.random(n) < 0.9).astype(np.uint8), prec=8)
        got = mt.oracle_grid_v_bound(sc)
        ok = sc["valid"] == 1
        # dx = (1/40, 1/40, 1/20): the bound scales |B| by 6 sqrt(d) times the
        # smallest 1/dx (MPM_DRIVER.cpp:1130), 20 in 3D and 40 in 2D
        xi = 6 * np.sqrt(dim) * (20.0 if dim == 3 else 40.0)
        want = (np.linalg.norm(sc["V"][ok], axis=1) + np.linalg.norm(sc["B"][ok], axis=1) * xi).max()
        assert abs(got - want) <= 1e-14 * want
        # the largest particle is excluded once it is invalid
        j = np.flatnonzero(ok)[np.argmax((np.linalg.norm(sc["V"], axis=1) + np.linalg.norm(sc["B"], axis=1) * xi)[ok])]
        sc["valid"] = sc["valid"].copy()
        sc["valid"][j] = 0
        assert mt.oracle_grid_v_bound(sc) < got


@pytest.mark.parametrize("name", ["snow2d_f64", "snow3d_f32", "sand2d_f64", "sand3d_f32"])
def test_oracle_plasticity_matches_reference(name):
    """MPM_PLASTICITY_CLAMP (snow) and MPM_DRUCKER_PRAGER on the Hencky-strain
    model (sand) after every step: Fp, the hardened mu and lambda or the
    Drucker-Prager state, and the projected F are bit-exact with the
    reference's own run."""
    z = load("plastic.npz")
    meta = json.loads(str(z[f"{name}_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[f"{name}_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    steps = sorted({int(k.split("_s")[1].split("_")[0]) for k in z.files if k.startswith(name + "_s")})
    assert steps
    sand = sc["plasticity"].get("kind") == "drucker_prager"
    keys = ("X", "V", "F", "B", "valid", "Fp", "mu", "lambda")
    if sand:
        keys += ("plastic_def", "rho_F", "cohesion", "vc")
    for n in steps:
        out = mt.oracle_run_drucker_prager(sc, n) if sand else mt.oracle_run_plastic(sc, n)
        for k in keys:
            got = out[k].astype(np.float64) if k != "valid" else out[k]
            assert np.array_equal(got, z[f"{name}_s{n}_{k}"]), (name, n, k)
    if sand:  # projected particles: plastic deformation accumulated, friction hardened
        assert (out["plastic_def"] > 0).mean() > 0.3
        assert out["rho_F"].max() > out["rho_F"].min()
    else:     # the fixture clamps: the Lame parameters hardened or softened
        assert (np.abs(out["mu"] - sc["mu"]) > 0).mean() > 0.5


def test_plasticity_leaves_elastic_range_alone():
    """Singular values inside [1 - theta_c, 1 + theta_s]: no change to F, Fp,
    mu or lambda (the clamp is the identity there)."""
    sc = mt.make_scene(2, 16, prec=8, shape=lambda X: np.abs(X[:, 0] - 0.5) < 0.2, gravity=[0.0, 0.0], dt=1e-5,
                       frand=1e-4)
    sc["plasticity"] = dict(theta_c=0.1, theta_s=0.1, max_hardening=5.0, hardening_factor=10.0)
    out = mt.oracle_run_plastic(sc, 1)
    ref = mt.oracle_run(sc, 1, want_grid=False)
    assert np.array_equal(out["F"], ref["F"])
    assert np.array_equal(out["Fp"], np.tile(np.eye(2).reshape(-1), (sc["X"].shape[0], 1)))
    assert np.array_equal(out["mu"], sc["mu"])


def test_plasticity_projection_closed_form():
    """After the update, the singular values of F lie in the clamp range,
    F Fp is the elastic-plastic total (F_before Fp_before), and the hardening
    factor is exp(xi (1 - det Fp)) clipped at exp(max_hardening)."""
    sc = mt.make_scene(3, 8, prec=8, shape=lambda X: np.abs(X[:, 0] - 0.5) < 0.3, gravity=[0.0, 0.0, 0.0],
                       dt=1e-6, frand=0.3, seed=5)
    pl = dict(theta_c=0.05, theta_s=0.02, max_hardening=0.5, hardening_factor=10.0)
    sc["plasticity"] = pl
    before = mt.oracle_run(sc, 1, want_grid=False)   # the elastic step alone
    out = mt.oracle_run_plastic(sc, 1)
    F = out["F"].reshape(-1, 3, 3).transpose(0, 2, 1)
    Fp = out["Fp"].reshape(-1, 3, 3).transpose(0, 2, 1)
    Fe0 = before["F"].reshape(-1, 3, 3).transpose(0, 2, 1)
    s = np.linalg.svd(F, compute_uv=False)
    assert (s >= 1 - pl["theta_c"] - 1e-12).all() and (s <= 1 + pl["theta_s"] + 1e-12).all()
    assert np.abs(F @ Fp - Fe0).max() < 1e-12
    h = np.exp(np.minimum(pl["max_hardening"], pl["hardening_factor"] * (1 - np.linalg.det(Fp))))
    assert np.allclose(out["mu"], sc["mu"] * h, rtol=1e-12)


LIMIT_CASES = ["strong2d_f64", "both2d_hencky_fric_f64", "sound3d_coro_f64", "sound3d_neo_f32", "strong3d_hencky_f32",
               "frames2d_f64"]


def limit_scene(z, name):
    meta = json.loads(str(z[f"{name}_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[f"{name}_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    return sc, meta


@pytest.mark.parametrize("name", LIMIT_CASES)
def test_oracle_dt_limits_match_reference(name):
    """Reduce_Dt (strong CFL, sound-speed CFL, min_dt, the 0.99 rule and the
    grid rescale) inside fixed-dt steps and inside the frame loop: the dts
    taken and the final state agree bit for bit with the reference's own runs
    (tests/golden/limits.npz)."""
    z = load("limits.npz")
    sc, m = limit_scene(z, name)
    run, lim = m["run"], m["limits"]
    if run[0] == "steps":
        out, dts = mt.oracle_run_limited(sc, run[1], lim)
    else:
        out, dts = mt.oracle_frames(sc, *run[1:], lim=lim)
    ref = z[f"{name}_dts"]
    assert np.array_equal(dts, ref), (dts, ref)
    for k in ("X", "V", "F", "B", "valid"):
        assert np.array_equal(out[k].astype(np.float64) if k != "valid" else out[k], z[f"{name}_out_{k}"]), k


def test_dt_limit_fixtures_lower_some_steps_and_keep_others():
    """The fixtures cover both outcomes of Reduce_Dt: steps whose dt is lowered
    (below 0.99 of the request) and steps that keep it."""
    z = load("limits.npz")
    lowered = kept = 0
    for name in LIMIT_CASES[:-1]:
        sc, m = limit_scene(z, name)
        r = z[f"{name}_dts"] / np.float64(np.float32(sc["dt"]) if sc["prec"] == 4 else sc["dt"])
        lowered += int((r < 0.99).sum())
        kept += int((r == 1).sum())
    assert lowered >= 10 and kept >= 5


@pytest.mark.parametrize("name", ["sound3d_coro_f64", "sound3d_neo_f32"])
def test_sound_speed_sets_the_first_dt_of_the_reference(name):
    """With only the sound-speed CFL on, the reference's first dt is
    cfl_sound * min dx / max_p c_p over the step's particles: the oracle's
    per-particle Sound_Speed reproduces it bit for bit."""
    z = load("limits.npz")
    sc, m = limit_scene(z, name)
    T = np.float32 if sc["prec"] == 4 else np.float64
    c = mt.oracle_sound_speeds(sc)[sc["valid"] == 1]
    dx = min(T(T(h) - T(l)) / T(n) for l, h, n in zip(sc["lo"], sc["hi"], sc["N"]))
    want = T(dx / c.max()) * T(m["limits"]["cfl_sound"])
    assert float(want) == z[f"{name}_dts"][0]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("model", [mt.COROTATED, mt.NEO_HOOKEAN, mt.HENCKY])
def test_sound_speed_at_rest_is_the_p_wave_speed(dim, model):
    """At F = I every model's stiffness is lambda + 2 mu along the axes (and mu
    in shear), so the sound speed is the P-wave speed sqrt((lambda + 2 mu) / rho);
    rotations leave it unchanged (it depends on the singular values only)."""
    rng = np.random.default_rng(dim * 10 + model)
    n = 50
    E = rng.uniform(1e3, 1e6, n)
    nu = rng.uniform(0.1, 0.45, n)
    mu, la = mt.lame(E, nu)
    rho = rng.uniform(500, 3000, n)
    vol = np.full(n, 1e-4)
    sc = dict(dim=dim, model=model, prec=8, X=np.zeros((n, dim)), mass=rho * vol, volume=vol, mu=mu, **{"lambda": la},
              F=np.tile(np.eye(dim).reshape(-1), (n, 1)))
    want = np.sqrt((la + 2 * mu) / rho)
    assert np.allclose(mt.oracle_sound_speeds(sc), want, rtol=1e-14)
    Q = np.linalg.qr(rng.normal(size=(n, dim, dim)))[0]
    Q[np.linalg.det(Q) < 0, :, 0] *= -1
    Fr = Q.transpose(0, 2, 1).reshape(n, -1)  # column-major storage of Q
    assert np.allclose(mt.oracle_sound_speeds(sc, F=Fr), want, rtol=1e-12)


def test_sound_speed_grows_under_compression_for_neo_hookean():
    """Neo-Hookean stiffens under compression: H_ii = mu + (lambda + mu - lambda
    ln J) / s_i^2 exceeds its rest value for J < 1, so the speed rises."""
    n = 3
    mu, la = mt.lame(np.full(n, 1e5), np.full(n, 0.3))
    sc = dict(dim=3, model=mt.NEO_HOOKEAN, prec=8, X=np.zeros((n, 3)), mass=np.full(n, 1.0), volume=np.full(n, 1e-3),
              mu=mu, **{"lambda": la}, F=np.array([np.diag([s] * 3).reshape(-1) for s in (1.0, 0.9, 0.8)]))
    c = mt.oracle_sound_speeds(sc)
    assert c[0] < c[1] < c[2]
    s = 0.8
    H = mu[2] + (la[2] + mu[2] - la[2] * np.log(s ** 3)) / s ** 2
    assert np.isclose(c[2], np.sqrt(H * s ** 2 / 1000.0), rtol=1e-13)


REFLECT_CASES = ["reflect2d_f64", "reflect2d_full_f64", "reflect3d_f32", "reflect3d_full_f64"]


@pytest.mark.parametrize("name", REFLECT_CASES)
def test_oracle_reflection_bc_matches_reference(name):
    """Reflection boundary conditions (mass-momentum, force and velocity
    reflection, or the frictional full variant; particles mirrored back at
    sides with a condition, invalidated at free ones): bit for bit with the
    reference's own runs after one step and after several
    (tests/golden/reflect.npz)."""
    z = load("reflect.npz")
    sc, m = limit_scene(z, name)
    sc["reflection"] = m["reflection"]
    for n in (1, m["nsteps"]):
        out = mt.oracle_run_reflect(sc, n)
        for k in ("X", "V", "F", "B", "valid"):
            assert np.array_equal(out[k].astype(np.float64) if k != "valid" else out[k], z[f"{name}_s{n}_{k}"]), (n, k)


def test_reflection_keeps_particles_that_free_sides_lose():
    """Particles leaving through a side with a condition come back mirrored
    (X -> 2 A - X, V_a -> -V_a); with every side free the same run loses them."""
    z = load("reflect.npz")
    sc, m = limit_scene(z, "reflect3d_f32")
    sc["reflection"] = m["reflection"]
    kept = mt.oracle_run_reflect(sc, m["nsteps"])
    free = dict(sc, reflection=dict(m["reflection"], sides=[mt.FREE] * 6))
    lost = mt.oracle_run_reflect(free, m["nsteps"])
    assert kept["valid"].sum() > lost["valid"].sum()
    X = kept["X"].astype(np.float64)
    assert np.all(X[kept["valid"] == 1] >= 0) and np.all(X[kept["valid"] == 1] <= 1)


# ------------------------------------------------------ conservation totals ----
TOTALS = ["rot3d_coro_f64", "rot2d_neo_f64", "rot3d_neo_floor_f32", "rot2d_hencky_f64"]


def totals_scene(z, name):
    meta = json.loads(str(z[name + "_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[name + "_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    return sc, meta["steps"]


@pytest.mark.parametrize("name", TOTALS)
def test_oracle_totals_match_reference(name):
    """Linear and angular momentum, kinetic and potential energy of particles
    and grid, as the reference's own MPM_EXAMPLE::Total_* and Potential_Energy
    report them around each step (recorded through its named-step callbacks),
    bit for bit."""
    z = load("totals.npz")
    sc, steps = totals_scene(z, name)
    ref = z[name + "_totals"]
    got = mt.oracle_totals(sc, steps)
    assert ref.shape == got.shape == (3 * steps + 1, 11)
    np.testing.assert_array_equal(got, ref)


def test_totals_fixtures_exercise_every_term():
    """The fixtures hold nonzero angular momentum, a B term, elastic energy of
    three models and gravity energy, so a dropped term changes a value."""
    z = load("totals.npz")
    for name in TOTALS:
        sc, _ = totals_scene(z, name)
        ref = z[name + "_totals"]
        assert np.abs(sc["B"]).max() > 0
        assert np.abs(ref[:, 4:7]).max() > 0
        assert ref[-1, 9] > 0, name  # elastic energy after the body deformed
    assert {totals_scene(z, n)[0]["model"] for n in TOTALS} == {0, 1, 2}
    assert abs(z["rot3d_neo_floor_f32_totals"][0, 10]) > 0 and abs(z["rot2d_hencky_f64_totals"][0, 10]) > 0


@pytest.mark.parametrize("dim,model,omega,N", [(3, mt.COROTATED, [0.4, 0.3, -0.2], 16), (2, mt.NEO_HOOKEAN, [3.0], 32),
                                              (3, mt.NEO_HOOKEAN, [2.0, -1.0, 5.0], 16)])
def test_apic_conserves_angular_momentum(dim, model, omega, N):
    """APIC's defining property: with no gravity and no walls, the angular
    momentum of the particles (with the -m B:eps term), of the grid after P2G
    and of the grid after the forces stays the same through every step (FP64
    rounding), and so do mass and linear momentum.  Dropping the B term, or
    using another node location or B update, breaks it at the 1e-4 level."""
    sc = mt.make_rotating(dim, N, prec=8, model=model, omega=omega, dt=2e-3, seed=41)
    o = mt.oracle_totals(sc, 20)
    L = np.abs(o[0, 4:7]).max()
    assert L > 0
    assert np.abs(o[:, 4:7] - o[0, 4:7]).max() <= 1e-13 * L
    assert np.abs(o[:, 0] - o[0, 0]).max() <= 1e-13 * o[0, 0]
    assert np.abs(o[:, 1:4]).max() <= 1e-13 * o[0, 0] * np.abs(sc["V"]).max()
    # without the -m B:eps term the particle total is off by far more
    m = sc["mass"]
    X, V = sc["X"], sc["V"]
    naive = (m[:, None] * (np.cross(X, V) if dim == 3 else (X[:, 0] * V[:, 1] - X[:, 1] * V[:, 0])[:, None])).sum(0)
    assert np.abs(naive[: (3 if dim == 3 else 1)] - o[0, 4:4 + (3 if dim == 3 else 1)]).max() > 1e-4 * L


def test_rigid_rotation_kinetic_energy_is_transferred_exactly():
    """A rigid rotation is an affine velocity field, which APIC transfers
    exactly: the particle kinetic energy with its B term equals the grid's
    after P2G."""
    sc = mt.make_rotating(3, 16, prec=8, model=mt.COROTATED, omega=[0.4, 0.3, -0.2], dt=2e-3, seed=41)
    o = mt.oracle_totals(sc, 1)
    assert abs(o[0, 7] - o[1, 7]) <= 1e-13 * o[0, 7]
