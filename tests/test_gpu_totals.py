"""Conservation totals on the GPU (mpmx_particle_totals, mpmx_grid_totals)
against the oracle, and APIC's angular-momentum conservation through the fused
and the unfused step.

The GPU sums in FP64 whatever the context precision, so an FP32 context is
held to the FP64 oracle run on the same float inputs.  Each quantity's error
is max|a - b| over its components divided by the largest magnitude of that
quantity (for the linear momentum, whose total is ~0 in the rotating scenes,
by M max|v|).
"""
import json
import os

import numpy as np
import pytest

import mpmtest as mt
from physbam_amd import mpmx
from test_gpu_step import gpu_context

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["rot3d_coro_f64", "rot2d_neo_f64", "rot3d_neo_floor_f32", "rot2d_hencky_f64"]
GROUPS = {"mass": slice(0, 1), "linear": slice(1, 4), "angular": slice(4, 7), "kinetic": slice(7, 8),
          "elastic": slice(9, 10), "gravity": slice(10, 11)}


def fixture_scene(name):
    z = np.load(os.path.join(GOLD, "totals.npz"), allow_pickle=False)
    meta = json.loads(str(z[name + "_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[name + "_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    if sc["prec"] == 4:  # an FP32 context holds gravity as a float: the FP64 oracle takes the same value
        sc["gravity"] = [float(np.float32(g)) for g in sc["gravity"]]
    return sc, meta["steps"]


def errors(got, ref, sc, groups=GROUPS):
    vmax = np.abs(sc["V"]).max()
    out = {}
    for k, s in groups.items():
        scale = np.abs(ref[s]).max()
        if k == "linear":
            scale = max(scale, ref[0] * vmax)
        if scale == 0:
            scale = 1.0
        out[k] = float(np.abs(np.asarray(got[s]) - np.asarray(ref[s])).max() / scale)
    return out


def check(tag, errs, tol):
    print(tag, {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, (tag, k, v)


@pytest.mark.parametrize("name", NAMES)
def test_particle_totals_match_oracle(name):
    """mpmx_particle_totals of the fixture's initial state against the oracle's
    restatement of Total_Particle_* and Potential_Energy (FP64, same inputs)."""
    sc, _ = fixture_scene(name)
    ctx = gpu_context(sc)
    got = ctx.particle_totals()
    ctx.close()
    ref = mt.oracle_totals(sc, 0, prec=8)[0]
    assert got[11] == int(np.count_nonzero(sc["valid"]))
    assert got[8] == pytest.approx(got[9] + got[10], rel=1e-15, abs=1e-300)
    check(name, errors(got, ref, sc), 1e-11)


@pytest.mark.parametrize("name", NAMES)
def test_grid_totals_match_oracle(name):
    """mpmx_grid_totals after P2G and after the grid update of each step against
    the oracle's Total_Grid_* of the same step: FP64 runs over all the
    fixture's steps (<=1e-9), FP32 runs after one step (<=1e-5; the FP32 node
    values themselves carry the step's rounding)."""
    sc, steps = fixture_scene(name)
    nsteps = steps if sc["prec"] == 8 else 1
    ref = mt.oracle_totals(sc, nsteps, prec=8)
    tol = 1e-9 if sc["prec"] == 8 else 1e-5
    ctx = gpu_context(sc)
    ctx.enable_grid_stats()
    grid_groups = {k: v for k, v in GROUPS.items() if k in ("mass", "linear", "angular", "kinetic")}
    for k in range(nsteps):
        ctx.step(sc["dt"], 1)
        for which in (0, 1):
            got = ctx.grid_totals(which)
            r = ref[3 * k + 1 + which]
            check(f"{name} step {k} grid {which}", errors(got, r, sc, grid_groups), tol)
            assert got[8] == 0 and got[9] == 0 and got[10] == 0
    ctx.close()


ROTATING = [  # dim, prec, model, omega, cells
    (3, 8, mt.COROTATED, [0.4, 0.3, -0.2], 16),
    (3, 4, mt.COROTATED, [0.4, 0.3, -0.2], 16),
    (2, 8, mt.NEO_HOOKEAN, [3.0], 32),
    (2, 4, mt.NEO_HOOKEAN, [3.0], 32),
    (3, 4, mt.NEO_HOOKEAN, [2.0, -1.0, 5.0], 32),
]


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("dim,prec,model,omega,N", ROTATING)
def test_rotating_body_conserves_angular_momentum(dim, prec, model, omega, N, fused):
    """The reference's rotating-ball scenes (no gravity, no walls) over 40
    steps: the particle angular momentum (with its -m B:eps term), mass and
    linear momentum stay constant -- to 1e-12 relative in FP64 and to FP32
    rounding (1e-5) in FP32.  Through the fused step (the particle totals after
    every step; the phase timers show the fused kernel ran) and through the
    unfused kernels, which every step takes while grid stats are enabled (then
    also the grid's totals after P2G and after the update)."""
    sc = mt.make_rotating(dim, N, prec=prec, model=model, omega=omega, dt=2e-3, seed=41)
    ctx = gpu_context(sc)
    if fused:
        ctx.enable_timing(True)
    else:
        ctx.enable_grid_stats()
    t0 = ctx.particle_totals()
    tol = 1e-12 if prec == 8 else 1e-5
    tol_m = 1e-12 if prec == 8 else 1e-6  # FP32 node masses are float sums
    ns = 3 if dim == 3 else 1
    L = np.abs(t0[4:4 + ns]).max()
    M = t0[0]
    P = M * np.abs(sc["V"]).max()
    assert L > 0
    worst = 0.0
    for k in range(40):
        ctx.step(sc["dt"], 1)
        tots = [ctx.particle_totals()]
        if not fused:
            tots += [ctx.grid_totals(0), ctx.grid_totals(1)]
        for t in tots:
            worst = max(worst, np.abs(t[4:4 + ns] - t0[4:4 + ns]).max() / L)
            assert abs(t[0] - M) <= tol_m * M
            assert np.abs(t[1:1 + dim]).max() <= tol * P
    t1 = ctx.particle_totals()
    worst = max(worst, np.abs(t1[4:4 + ns] - t0[4:4 + ns]).max() / L)
    if fused:
        ctx.synchronize()
        assert ctx.phase_time("g2p2g")[1] >= 39  # every step after the first ran the fused kernel
        assert ctx.phase_time("g2p")[1] == 0
    ctx.close()
    print(f"rotating d{dim} p{prec} m{model} fused={fused}: angular momentum drift {worst:.2e}", flush=True)
    assert abs(t1[0] - M) <= 1e-12 * M
    assert t1[11] == t0[11]
    assert worst <= tol


def test_totals_follow_the_state():
    """Particle totals change with the state they are taken of: after a
    constant velocity is added the linear momentum is M dv, the kinetic
    energy grows by M |dv|^2 / 2 (the rotation's momentum is zero), and the
    gravity energy is -M x_c . g of the centre of mass."""
    sc = mt.make_rotating(3, 16, prec=8, model=mt.COROTATED, omega=[0.0, 0.0, 1.0], dt=1e-3, seed=3,
                          gravity=[0.0, -9.8, 0.0])
    ctx = gpu_context(sc)
    t0 = ctx.particle_totals()
    dv = np.array([0.5, -0.25, 1.0])
    sc2 = dict(sc)
    sc2["V"] = sc["V"] + dv
    ctx.close()
    ctx = gpu_context(sc2)
    t1 = ctx.particle_totals()
    ctx.close()
    M = t0[0]
    np.testing.assert_allclose(t1[1:4], M * dv, rtol=1e-12, atol=1e-12 * M)
    assert t1[7] - t0[7] == pytest.approx(0.5 * M * (dv @ dv), rel=1e-10)
    xc = (sc["mass"][:, None] * sc["X"]).sum(0) / M
    assert t0[10] == pytest.approx(-M * (xc @ np.array([0.0, -9.8, 0.0])), rel=1e-12)
    assert t0[9] == pytest.approx(0.0, abs=1e-20)  # F = I: no elastic energy
