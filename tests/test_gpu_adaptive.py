"""Adaptive time stepping on the GPU (mpmx_grid_v_upper_bound,
mpmx_compute_dt, mpmx_advance_frames) against the reference's substep loop
(tests/golden/frames.npz, written by the reference's own Compute_Dt and
Simulate_To_Frame loop) and the oracle.

- speed bound: bit-identical to the oracle's Grid_V_Upper_Bound for the same
  particle state (FP32 and FP64, 2D and 3D, with invalid particles), per
  scene in a batch, and at 4M particles to the same formula evaluated in the
  context precision by numpy (IEEE operations in the same order).
- compute_dt: equal to Compute_Dt formed in the context precision from that
  bound, on both branches (CFL-limited and max_dt).
- frames: the same number of substeps as the reference; FP64 dts to 1e-9
  relative and the final state to 1e-7; the FP32 scene's dts to 1e-3 and X to
  1e-5, V and F to 1e-4 (the step itself is compared per step elsewhere).
"""
import json
import os

import numpy as np
import pytest

import mpmtest as mt
from physbam_amd import mpmx
from test_gpu_step import gpu_context, rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def frame_fixture(name):
    z = np.load(os.path.join(GOLD, "frames.npz"), allow_pickle=False)
    meta = json.loads(str(z[f"{name}_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[f"{name}_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    return z, sc, meta


def bound_scene(dim, prec, seed, n_invalid=50, scale=1.0):
    sph = (lambda X: ((X - 0.5) ** 2).sum(1) < 0.3 ** 2)
    sc = mt.make_scene(dim, [24, 16, 20][:dim], prec=prec, shape=sph, v0=[1.0, -2.0, 0.5][:dim], vrand=2.0,
                       brand=0.3, frand=0.1, seed=seed, hi=[1.0, 0.8, 1.2][:dim])
    sc["V"] = sc["V"] * scale
    if prec == 4:
        sc["V"] = sc["V"].astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(seed)
    sc["valid"] = sc["valid"].copy()
    sc["valid"][rng.choice(sc["X"].shape[0], n_invalid, replace=False)] = 0
    # the fastest particle is invalid: the bound must not see it
    j = int(np.argmax(np.linalg.norm(sc["V"], axis=1)))
    sc["valid"][j] = 0
    return sc


def expected_dt(sc, v, cfl, max_dt):
    """Compute_Dt (MPM_DRIVER.cpp:1098-1104) in the scene's precision."""
    T = np.float32 if sc["prec"] == 4 else np.float64
    dx = [(T(h) - T(l)) / T(n) for l, h, n in zip(sc["lo"], sc["hi"], sc["N"])]
    dxmin = min(dx)
    crit = T(cfl) * dxmin / T(max_dt)
    v = T(v)
    return float(T(cfl) * dxmin / v if v > crit else T(max_dt))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", [4, 8])
def test_speed_bound_is_bit_identical_to_oracle(dim, prec):
    sc = bound_scene(dim, prec, seed=10 * dim + prec)
    print("bound", dim, prec, sc["X"].shape[0], flush=True)
    want = mt.oracle_grid_v_bound(sc)
    with gpu_context(sc) as ctx:
        got = ctx.grid_v_upper_bound()
        assert got.shape == (1,)
        assert got[0] == want, (got[0], want)
        for cfl, max_dt, scale in ((0.5, 1e-2, 1.0), (0.9, 1e-4, 1.0)):
            dt = ctx.compute_dt(cfl, max_dt)[0]
            assert dt == expected_dt(sc, want, cfl, max_dt), (dt, cfl, max_dt)
        # both branches are taken above: CFL-limited, then max_dt
        assert expected_dt(sc, want, 0.5, 1e-2) < 1e-2
        assert expected_dt(sc, want, 0.9, 1e-4) == pytest.approx(1e-4, rel=1e-6)


def test_speed_bound_per_scene_in_a_batch():
    scs = [bound_scene(3, 4, seed=s, scale=sc_) for s, sc_ in ((1, 1.0), (2, 3.0), (3, 0.2))]
    scs[2]["valid"][:] = 0  # a scene with no valid particle: bound 0, dt = max_dt
    ctx = mpmx.Context.batch(3, [(s["N"], s["lo"], s["hi"]) for s in scs], precision=mpmx.FP32)
    try:
        cat = lambda k: np.concatenate([s[k] for s in scs])
        ctx.set_batch_particles([s["X"].shape[0] for s in scs], cat("X"), cat("V"), cat("F"), cat("B"), cat("mass"),
                                cat("volume"), cat("mu"), cat("lambda"), cat("valid"))
        got = ctx.grid_v_upper_bound()
        want = [mt.oracle_grid_v_bound(s) for s in scs]
        print("batch bounds", got, want, flush=True)
        assert list(got) == want
        assert want[2] == 0.0
        dts = ctx.compute_dt(0.5, 1e-3)
        assert list(dts) == [expected_dt(s, w, 0.5, 1e-3) for s, w in zip(scs, want)]
        with pytest.raises(mpmx.MpmxError):
            ctx.advance_frames(1, 1e-3)
    finally:
        ctx.close()


def test_speed_bound_at_4m_particles_matches_numpy_in_fp32():
    rng = np.random.default_rng(77)
    n, d = 4_000_000, 3
    N, lo, hi = [128, 96, 160], [0.0, 0.0, 0.0], [1.0, 0.75, 1.0]
    X = (rng.random((n, d)) * 0.8 + 0.1) * np.array(hi)
    V = (rng.standard_normal((n, d)) * 2).astype(np.float32)
    B = (rng.standard_normal((n, d * d)) * 1e-2).astype(np.float32)
    valid = (rng.random(n) < 0.97).astype(np.uint8)
    F = np.tile(np.eye(d, dtype=np.float32).reshape(-1), (n, 1))
    ones = np.ones(n, np.float32)
    with mpmx.Context(d, N, lo, hi, precision=mpmx.FP32) as ctx:
        ctx.set_particles(X, V, F, B, ones * 1e-3, ones * 1e-6, ones * 1e3, ones * 1e3, valid)
        got = ctx.grid_v_upper_bound()[0]
    f = np.float32
    v2 = V[:, 0] * V[:, 0]
    for c in range(1, d):
        v2 = v2 + V[:, c] * V[:, c]
    b2 = B[:, 0] * B[:, 0]
    for k in range(1, d * d):
        b2 = b2 + B[:, k] * B[:, k]
    odx = min(f(1) / ((f(h) - f(l)) / f(m)) for l, h, m in zip(lo, hi, N))
    xi = f(6) * np.sqrt(f(d)) * odx
    r = np.sqrt(v2) + np.sqrt(b2) * xi
    want = float(r[valid == 1].max())
    print("4M bound", got, want, flush=True)
    assert got == want


@pytest.mark.parametrize("name", ["throw2d_f64", "ball3d_neo_slip_f32"])
def test_adaptive_frames_match_reference(name):
    z, sc, m = frame_fixture(name)
    fp32 = sc["prec"] == 4
    with gpu_context(sc) as ctx:
        t0 = ctx.time
        dts = ctx.advance_frames(m["nframes"], m["frame_dt"], m["cfl"], m["min_dt"], m["max_dt"])
        got = ctx.get_particles()
        t1 = ctx.time
    ref = z[f"{name}_dts"]
    print(name, "substeps", len(dts), "reference", len(ref), flush=True)
    assert len(dts) == len(ref)
    e = float(np.abs(dts - ref).max() / ref.max())
    print(f"  dts {e:.2e}", flush=True)
    assert e <= (1e-3 if fp32 else 1e-9)
    assert t1 - t0 == pytest.approx(m["nframes"] * m["frame_dt"], rel=1e-6)
    assert np.array_equal(got["valid"], z[f"{name}_out_valid"])
    tols = dict(X=1e-5, V=1e-4, F=1e-4) if fp32 else dict(X=1e-7, V=1e-7, F=1e-7)
    for k, tol in tols.items():
        err = rel(got[k], z[f"{name}_out_{k}"])
        print(f"  {k} {err:.2e} (tol {tol:.0e})", flush=True)
        assert err <= tol, (k, err)


def test_adaptive_frames_rejects_bad_arguments():
    _, sc, m = frame_fixture("throw2d_f64")
    with gpu_context(sc) as ctx:
        for args in ((1, -1.0, 0.5, 0.0, 1e-3), (1, 1e-2, 0.0, 0.0, 1e-3), (1, 1e-2, 0.5, 2e-3, 1e-3)):
            with pytest.raises(mpmx.MpmxError):
                ctx.advance_frames(*args)
        assert ctx.time == 0.0
        # a frame shorter than max_dt: one substep landing on the frame end
        dts = ctx.advance_frames(1, 1e-6, 0.5, 0.0, 1e-3)
        assert list(dts) == [1e-6]
