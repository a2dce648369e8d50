"""Clamped plasticity with hardening on the GPU (mpmx_set_plasticity) against
the reference's own runs (tests/golden/plastic.npz, MPM_PLASTICITY_CLAMP after
every step) and the oracle.

FP64: X, V, F, B, Fp, mu, lambda to 1e-9 relative after one step and 1e-8
after ten.  FP32: against the FP64 oracle of the same float inputs, X, V, F,
Fp, mu, lambda to 1e-5 and B to the bound the golden step test uses (B is
measured against max(|B|, |V| h)).  The clamp is continuous in F, so the
plastic update keeps these tolerances.
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


def fixture(name):
    z = np.load(os.path.join(GOLD, "plastic.npz"), allow_pickle=False)
    sc = json.loads(str(z[f"{name}_meta"]))
    sc["walls"] = json.loads(str(z[f"{name}_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    steps = sorted({int(k.split("_s")[1].split("_")[0]) for k in z.files if k.startswith(name + "_s")})
    return z, sc, steps


def plastic_context(sc):
    ctx = gpu_context(sc)
    pl = sc["plasticity"]
    ctx.set_plasticity(pl["theta_c"], pl["theta_s"], pl["max_hardening"], pl["hardening_factor"])
    return ctx


@pytest.mark.parametrize("name", ["snow2d_f64", "snow3d_f32"])
def test_plasticity_matches_reference(name):
    z, sc, steps = fixture(name)
    fp32 = sc["prec"] == 4
    h = (np.asarray(sc["hi"]) - np.asarray(sc["lo"])).max() / max(sc["N"])
    with plastic_context(sc) as ctx:
        done = 0
        for n in steps:
            print(name, "steps", n, flush=True)
            ctx.step(sc["dt"], n - done)
            done = n
            got = dict(ctx.get_particles())
            got.update(ctx.get_plastic_state())
            if fp32:  # the exact answer for the float inputs: the FP64 oracle
                ref = mt.oracle_run_plastic(sc, n, prec=8)
            else:
                ref = {k: z[f"{name}_s{n}_{k}"] for k in ("X", "V", "F", "B", "valid", "Fp", "mu", "lambda")}
            assert np.array_equal(got["valid"], ref["valid"]), (name, n)
            for k in ("X", "V", "F", "B", "Fp", "mu", "lambda"):
                r = np.asarray(ref[k], np.float64)
                scale = max(np.abs(r).max(), np.abs(np.asarray(ref["V"])).max() * h) if k == "B" else np.abs(r).max()
                err = float(np.abs(got[k] - r).max() / scale)
                tol = (1e-5 if k != "B" else 3e-4 * max(1, n / 2)) if fp32 else (1e-9 if n <= 1 else 1e-8)
                print(f"  {k} {err:.2e} (tol {tol:.0e})", flush=True)
                assert err <= tol, (name, n, k, err)


def test_plasticity_off_again_is_the_elastic_step():
    _, sc, _ = fixture("snow2d_f64")
    with plastic_context(sc) as ctx:
        ctx.set_plasticity(0.02, 0.01, 5.0, 10.0, enable=False)
        ctx.step(sc["dt"], 3)
        got = ctx.get_particles()
        with pytest.raises(mpmx.MpmxError):
            ctx.get_plastic_state()
    orc = mt.oracle_run(sc, 3, want_grid=False)
    assert rel(got["F"], orc["F"]) < 1e-9


def test_plasticity_rejects_bad_parameters_and_tape():
    _, sc, _ = fixture("snow2d_f64")
    with gpu_context(sc) as ctx:
        for args in ((-0.1, 0.01, 5, 10), (1.0, 0.01, 5, 10), (0.02, -1.0, 5, 10), (0.02, 0.01, float("nan"), 10)):
            with pytest.raises(mpmx.MpmxError):
                ctx.set_plasticity(*args)
        ctx.set_plasticity(0.02, 0.01, 5.0, 10.0)
        with pytest.raises(mpmx.MpmxError):
            ctx.tape_begin(2)
