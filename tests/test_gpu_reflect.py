"""Reflection boundary conditions on the GPU (mpmx_set_reflection) against the
reference's own runs (tests/golden/reflect.npz) and the oracle.

FP64: X, V, F, B after one step and after the run to 1e-9 relative (B against
max(|B|, |V| h)), the valid flags exactly (particles mirrored back at sides
with a condition, invalidated at free ones).  FP32: against the FP64 oracle of
the same float inputs, X to 1e-5 and V, F, B to 1e-4.
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
CASES = ["reflect2d_f64", "reflect2d_full_f64", "reflect3d_f32", "reflect3d_full_f64"]


def fixture(name):
    z = np.load(os.path.join(GOLD, "reflect.npz"), allow_pickle=False)
    meta = json.loads(str(z[f"{name}_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[f"{name}_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    return z, sc, meta


def reflect_context(sc):
    ctx = gpu_context(sc)
    r = sc["reflection"]
    ctx.set_reflection(r["sides"], r.get("full", False), r.get("friction", 0.0))
    return ctx


@pytest.mark.parametrize("name", CASES)
def test_reflection_matches_reference(name):
    z, sc, m = fixture(name)
    fp32 = sc["prec"] == 4
    h = (np.asarray(sc["hi"]) - np.asarray(sc["lo"])).max() / max(sc["N"])
    with reflect_context(sc) as ctx:
        done = 0
        for n in (1, m["nsteps"]):
            print(name, "steps", n, flush=True)
            ctx.step(sc["dt"], n - done)
            done = n
            got = ctx.get_particles()
            if fp32:
                ref = mt.oracle_run_reflect(sc, n, prec=8)
            else:
                ref = {k: z[f"{name}_s{n}_{k}"] for k in ("X", "V", "F", "B", "valid")}
            assert np.array_equal(got["valid"], ref["valid"]), (name, n)
            for k in ("X", "V", "F", "B"):
                r = np.asarray(ref[k], np.float64)
                scale = max(np.abs(r).max(), np.abs(np.asarray(ref["V"])).max() * h) if k == "B" else np.abs(r).max()
                err = float(np.abs(got[k] - r).max() / scale)
                tol = ((1e-5 if k == "X" else 1e-4) if fp32 else 1e-9)
                print(f"  {k} {err:.2e} (tol {tol:.0e})", flush=True)
                assert err <= tol, (name, n, k, err)


def test_reflection_off_again_is_the_plain_step():
    _, sc, _ = fixture("reflect2d_f64")
    with reflect_context(sc) as ctx:
        ctx.set_reflection(sc["reflection"]["sides"], enable=False)
        ctx.step(sc["dt"], 2)
        got = ctx.get_particles()
    orc = mt.oracle_run(sc, 2, want_grid=False)
    assert np.array_equal(got["valid"], orc["valid"])
    assert rel(got["X"], orc["X"]) < 1e-12


def test_reflection_rejects_bad_arguments_and_combinations():
    _, sc, _ = fixture("reflect2d_f64")
    with gpu_context(sc) as ctx:
        for sides in ([0, 1, 2, 4], [-1, 0, 0, 0]):
            with pytest.raises(mpmx.MpmxError):
                ctx.set_reflection(sides)
        with pytest.raises(mpmx.MpmxError):
            ctx.set_reflection([1, 1, 1, 1], full=True, friction=-1.0)
        ctx.set_reflection([1, 1, 1, 1])
        with pytest.raises(mpmx.MpmxError):
            ctx.set_dt_limits(True, False)
        with pytest.raises(mpmx.MpmxError):
            ctx.tape_begin(2)
        ctx.set_reflection([1, 1, 1, 1], enable=False)
        ctx.set_dt_limits(True, False)
        with pytest.raises(mpmx.MpmxError):
            ctx.set_reflection([1, 1, 1, 1])
    small = mt.make_scene(2, 5, prec=8, dt=1e-4)  # fewer than 6 cells per axis
    with gpu_context(small) as ctx:
        with pytest.raises(mpmx.MpmxError):
            ctx.set_reflection([1, 1, 1, 1])
    # the full variant with a frictional collision object is refused at the step
    with gpu_context(sc) as ctx:
        ctx.add_collision_box(mpmx.SEPARATE, 0.3, [-1.0, -1.0], [1.0, 0.05])
        ctx.set_reflection([3, 3, 3, 3], full=True, friction=0.2)
        with pytest.raises(mpmx.MpmxError):
            ctx.step(sc["dt"], 1)
