"""Frame output and restart through a context (mpmx_write_output_files,
mpmx_read_output_files) against the reference's own files
(tests/golden/io.npz).

* Reading the reference's frame directory gives back, bit for bit, what the
  reference's own Read_From_File returns (particles, plastic state, time).
* Writing after the same steps gives the reference's attribute set and layout,
  and values equal to the context's state bit for bit; against the
  reference's written values FP64 agrees to 1e-9 and FP32 to 1e-5, B measured
  against max(|B|, |V| h) (the steps themselves are compared elsewhere).
* A run split by a write and a read continues as the unsplit run (FP64 to
  1e-12: only the particle storage order, and with it summation order, may
  differ).
"""
import gzip
import json
import os

import numpy as np
import pytest

import mpmtest as mt
from physbam_amd import mpmx
from test_gpu_step import gpu_context, rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["io2d_f64", "io3d_f32", "io2d_snow_f64"]
BASE = {"X": mpmx.ARRAY_VECTOR, "V": mpmx.ARRAY_VECTOR, "F": mpmx.ARRAY_MATRIX, "B": mpmx.ARRAY_MATRIX,
        "color": mpmx.ARRAY_VECTOR3, "mass": mpmx.ARRAY_SCALAR, "volume": mpmx.ARRAY_SCALAR,
        "valid": mpmx.ARRAY_BOOL, "mu": mpmx.ARRAY_SCALAR, "lambda": mpmx.ARRAY_SCALAR}
PLASTIC = {"Fp": mpmx.ARRAY_MATRIX, "mu0": mpmx.ARRAY_SCALAR, "lambda0": mpmx.ARRAY_SCALAR}


def fixture(name):
    z = np.load(os.path.join(GOLD, "io.npz"), allow_pickle=False)
    meta = json.loads(str(z[f"{name}_meta"]))
    sc = dict(meta)
    sc["walls"] = json.loads(str(z[f"{name}_walls"]))
    for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
        sc[k] = z[f"{name}_in_{k}"]
    return z, sc, meta


def context(sc):
    ctx = gpu_context(sc)
    if "plasticity" in sc:
        pl = sc["plasticity"]
        ctx.set_plasticity(pl["theta_c"], pl["theta_s"], pl["max_hardening"], pl["hardening_factor"])
    return ctx


def reference_dir(z, name, tmp_path):
    d = os.path.join(str(tmp_path), "ref")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "deformable_object_particles.gz"), "wb") as f:
        f.write(bytes(z[f"{name}_particles_gz"]))
    os.symlink("./deformable_object_particles.gz", os.path.join(d, "mpm_particles.gz"))
    with open(os.path.join(d, "restart_data.gz"), "wb") as f:
        f.write(bytes(z[f"{name}_restart_gz"]))
    return d


@pytest.mark.parametrize("name", CASES)
def test_read_output_files_matches_the_reference_reader(name, tmp_path):
    z, sc, m = fixture(name)
    d = reference_dir(z, name, tmp_path)
    print(name, flush=True)
    with context(sc) as ctx:
        ctx.read_output_files(d)
        got = dict(ctx.get_particles())
        got.update(ctx.get_material())
        T = np.float32 if sc["prec"] == 4 else np.float64
        assert ctx.time == float(T(z[f"{name}_read_time"]))
        if "plasticity" in sc:
            got.update(ctx.get_plastic_state())
        for k in ("X", "V", "F", "B", "valid", "mass", "volume", "mu", "lambda") + (("Fp",) if "plasticity" in sc else ()):
            want = z[f"{name}_read_{k}"]
            assert np.array_equal(np.asarray(got[k]).reshape(want.shape).astype(want.dtype), want), k


@pytest.mark.parametrize("name", CASES)
def test_write_output_files_after_steps(name, tmp_path):
    z, sc, m = fixture(name)
    d, prec = sc["dim"], sc["prec"]
    out = os.path.join(str(tmp_path), "frame")
    os.makedirs(out)
    print(name, flush=True)
    with context(sc) as ctx:
        ctx.step(sc["dt"], m["nsteps"])
        ctx.write_output_files(out)
        state = dict(ctx.get_particles())
        if "plasticity" in sc:
            state.update(ctx.get_plastic_state())
        t = ctx.time
    assert os.path.islink(os.path.join(out, "mpm_particles.gz"))
    kinds = dict(BASE, **(PLASTIC if "plasticity" in sc else {}))
    mine = mpmx.read_particles_file(os.path.join(out, "mpm_particles.gz"), d, prec, kinds)
    assert set(mine) == set(kinds)
    for k in ("X", "V", "F", "B", "valid") + (("Fp", "mu", "lambda") if "plasticity" in sc else ()):
        assert np.array_equal(mine[k].reshape(np.shape(state[k])), state[k]), k
    assert mpmx.read_restart_file(os.path.join(out, "restart_data.gz")) == t
    # the reference's file after the same steps: same arrays, same names and
    # types, values equal to the step's precision
    ref_raw = gzip.decompress(bytes(z[f"{name}_particles_gz"]))
    with open(os.path.join(out, "deformable_object_particles.gz"), "rb") as f:
        my_raw = gzip.decompress(f.read())
    assert len(my_raw) == len(ref_raw)
    assert my_raw[:13] == ref_raw[:13]  # stream type, version, count, array count
    ref = mpmx.read_particles_file(os.path.join(reference_dir(z, name, tmp_path), "mpm_particles.gz"), d, prec, kinds)
    assert np.array_equal(mine["valid"], ref["valid"])
    for k in ("mass", "volume", "color"):
        assert np.array_equal(mine[k], ref[k]), k
    h = (np.asarray(sc["hi"]) - np.asarray(sc["lo"])).max() / max(sc["N"])
    for k in ("X", "V", "F", "B") + (("Fp", "mu", "lambda") if "plasticity" in sc else ()):
        r = np.asarray(ref[k], np.float64)
        # B against max(|B|, |V| h), as the step tests measure it
        scale = max(np.abs(r).max(), np.abs(ref["V"]).max() * h) if k == "B" else np.abs(r).max()
        err = float(np.abs(mine[k] - r).max() / scale)
        print(f"  {k} {err:.2e}", flush=True)
        assert err <= (1e-5 if prec == 4 else 1e-9), (k, err)


def test_restart_continues_the_run(tmp_path):
    z, sc, m = fixture("io2d_f64")
    with context(sc) as ctx:
        ctx.step(sc["dt"], 6)
        whole = ctx.get_particles()
    out = str(tmp_path)
    with context(sc) as ctx:
        ctx.step(sc["dt"], 3)
        ctx.write_output_files(out)
    with context(sc) as ctx:
        ctx.read_output_files(out)
        assert ctx.time == pytest.approx(3 * sc["dt"], rel=1e-12)
        ctx.step(sc["dt"], 3)
        split = ctx.get_particles()
    assert np.array_equal(split["valid"], whole["valid"])
    for k in ("X", "V", "F", "B"):
        err = rel(split[k], whole[k])
        print(k, f"{err:.2e}", flush=True)
        assert err <= 1e-12, (k, err)


def test_output_files_reject_batches_and_missing_files(tmp_path):
    z, sc, m = fixture("io2d_f64")
    with context(sc) as ctx:
        with pytest.raises(mpmx.MpmxError):
            ctx.read_output_files(str(tmp_path))  # empty directory
        with pytest.raises(mpmx.MpmxError):
            ctx.write_output_files(os.path.join(str(tmp_path), "no", "such", "dir"))
    ctx = mpmx.Context.batch(2, [(sc["N"], sc["lo"], sc["hi"])] * 2, precision=mpmx.FP64)
    try:
        with pytest.raises(mpmx.MpmxError):
            ctx.write_output_files(str(tmp_path))
    finally:
        ctx.close()
