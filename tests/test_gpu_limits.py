"""Reduce_Dt on the GPU (mpmx_set_dt_limits: strong CFL, sound-speed CFL,
min_dt) against the reference's own runs (tests/golden/limits.npz) and the
oracle, and mpmx_max_sound_speed against the oracle's Sound_Speed.

FP64: the dts taken to 1e-12 relative and X, V, F, B to 1e-9 after the run.
FP32: against the FP64 oracle of the same float inputs, the dts to 1e-4
relative (the limits are formed from FP32 grid velocities and singular values)
and X to 1e-5, V, F, B to 1e-4 over the run; every fixture step is far from
the 0.99 rule's threshold, so both precisions take the same branch.
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
CASES = ["strong2d_f64", "both2d_hencky_fric_f64", "sound3d_coro_f64", "sound3d_neo_f32", "strong3d_hencky_f32",
         "frames2d_f64"]


def fixture(name):
    z = np.load(os.path.join(GOLD, "limits.npz"), allow_pickle=False)
    meta = json.loads(str(z[f"{name}_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[f"{name}_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    return z, sc, meta


def set_limits(ctx, lim):
    ctx.set_dt_limits(lim.get("strong_cfl", False), lim.get("sound_cfl", False), lim.get("cfl", 1.0),
                      lim.get("cfl_F", 0.5), lim.get("cfl_sound", 0.9), lim.get("min_dt", 0.0))


@pytest.mark.parametrize("name", CASES)
def test_dt_limits_match_reference(name):
    z, sc, m = fixture(name)
    fp32 = sc["prec"] == 4
    run, lim = m["run"], m["limits"]
    print(name, run, lim, flush=True)
    with gpu_context(sc) as ctx:
        set_limits(ctx, lim)
        t0 = ctx.time
        if run[0] == "steps":
            dts = []
            for _ in range(run[1]):
                ctx.step(sc["dt"], 1)
                dts.append(ctx.last_dt)
            dts = np.array(dts)
        else:
            dts = ctx.advance_frames(*run[1:])
        got = ctx.get_particles()
        assert ctx.time - t0 == pytest.approx(dts.sum(), rel=1e-6)
    if fp32:  # the exact answer for the float inputs: the FP64 oracle
        if run[0] == "steps":
            ref, rdts = mt.oracle_run_limited(sc, run[1], lim, prec=8)
        else:
            ref, rdts = mt.oracle_frames(sc, *run[1:], lim=lim, prec=8)
        ref = dict(ref)
    else:
        ref = {k: z[f"{name}_out_{k}"] for k in ("X", "V", "F", "B", "valid")}
        rdts = z[f"{name}_dts"]
    print("  dts", dts / sc["dt"], flush=True)
    assert len(dts) == len(rdts)
    e = float(np.abs(dts - rdts).max() / rdts.max())
    print(f"  dts {e:.2e}", flush=True)
    assert e <= (1e-4 if fp32 else 1e-12)
    assert np.array_equal(got["valid"], ref["valid"])
    for k in ("X", "V", "F", "B"):
        err = rel(got[k], ref[k])
        tol = (1e-5 if k == "X" else 1e-4) if fp32 else 1e-9
        print(f"  {k} {err:.2e} (tol {tol:.0e})", flush=True)
        assert err <= tol, (k, err)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("model", [mt.COROTATED, mt.NEO_HOOKEAN, mt.HENCKY])
def test_max_sound_speed_matches_oracle(dim, prec, model):
    sph = lambda X: ((X - 0.5) ** 2).sum(1) < 0.3 ** 2
    sc = mt.make_scene(dim, [24, 20, 16][:dim], prec=prec, model=model, shape=sph, frand=0.4, E=1e5,
                       per_particle_E=True, seed=dim * 100 + prec * 10 + model)
    sc["valid"] = sc["valid"].copy()
    c = mt.oracle_sound_speeds(sc, prec=8)
    sc["valid"][np.argmax(c)] = 0  # the fastest particle is invalid: the max must skip it
    if prec == 4:
        sc = dict(sc, **{k: np.asarray(sc[k], np.float32).astype(np.float64)
                         for k in ("F", "mass", "volume", "mu", "lambda")})
        c = mt.oracle_sound_speeds(sc, prec=8)
    want = c[sc["valid"] == 1].max()
    with gpu_context(sc) as ctx:
        got = ctx.max_sound_speed()
    print(dim, prec, model, got, want, flush=True)
    assert abs(got - want) <= (1e-5 if prec == 4 else 1e-12) * want


def test_dt_limits_reject_bad_arguments_batches_and_tapes():
    _, sc, _ = fixture("strong2d_f64")
    with gpu_context(sc) as ctx:
        for args in ((True, False, 0.0, 0.5, 0.9, 0.0), (True, False, 1.0, -1.0, 0.9, 0.0),
                     (False, True, 1.0, 0.5, 0.0, 0.0), (False, True, 1.0, 0.5, 0.9, -1.0)):
            with pytest.raises(mpmx.MpmxError):
                ctx.set_dt_limits(*args)
        ctx.set_dt_limits(True, False, 0.02, 0.1)
        with pytest.raises(mpmx.MpmxError):
            ctx.tape_begin(2)
        # both off: the step keeps dt and is the plain step
        ctx.set_dt_limits(False, False)
        ctx.step(sc["dt"], 2)
        assert ctx.last_dt == sc["dt"]
        got = ctx.get_particles()
    orc = mt.oracle_run(sc, 2, want_grid=False)
    assert rel(got["F"], orc["F"]) < 1e-9
    scs = [sc, sc]
    ctx = mpmx.Context.batch(2, [(s["N"], s["lo"], s["hi"]) for s in scs], precision=mpmx.FP64)
    try:
        with pytest.raises(mpmx.MpmxError):
            ctx.set_dt_limits(True, False)
        with pytest.raises(mpmx.MpmxError):
            ctx.max_sound_speed()
    finally:
        ctx.close()
