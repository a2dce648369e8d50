"""The drop-in inside PhysBAM itself: the reference's own MPM_EXAMPLE (set up
from a scene file exactly as the reference harness sets it up) advanced by
the adapter physbam_amd/host/MPM_GPU_STEP.h in place of
MPM_DRIVER::Advance_One_Time_Step (oracle/_ref/adapter_check), against the
reference's own CPU step of the same example (oracle/_ref/ref_harness step).
FP64 scenes, particle state after the steps and example.mass /
example.velocity read back from the device: 1e-9 relative (max|a - b| /
max|b| per quantity; B against max(|B|, |V| h)).  The same with begin/end
callbacks on every named step (the step then runs stage by stage).  Skipped
where the two programs were not built (they need the reference checkout)."""
import os
import subprocess

import numpy as np
import pytest

import mpmtest as mt
import scene_io

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "adapter_check")

sq2 = lambda X: (np.abs(X[:, 0] - 0.5) <= 0.2) & (np.abs(X[:, 1] - 0.6) <= 0.2)
sph3 = lambda X: ((X - np.array([0.5, 0.45, 0.5])) ** 2).sum(1) < 0.3 ** 2


def scenes():
    out = {}
    walls = mt.wall_boxes(2, [64, 64], [0, 0], [1, 1], flags=8, inset=0.05, type_=mt.STICK, friction=0.0)
    out["square2d"] = (mt.make_scene(2, 64, prec=8, model=mt.COROTATED, shape=sq2, v0=[0.5, -2.0], vrand=0.5,
                                     brand=1e-3, frand=0.1, walls=walls, dt=1e-4, seed=11), 20)
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.SEPARATE, friction=0.3)
    out["sphere3d_fric"] = (mt.make_scene(3, 16, prec=8, model=mt.NEO_HOOKEAN, shape=sph3, v0=[0.3, -3.0, 0.2],
                                          vrand=1.0, brand=1e-3, frand=0.2, walls=walls, dt=1e-3,
                                          per_particle_E=True, seed=12), 8)
    return out


@pytest.mark.skipif(not (os.access(EXE, os.X_OK) and scene_io.ref_available()),
                    reason="oracle/_ref/adapter_check and ref_harness are built only with the reference checkout")
@pytest.mark.parametrize("stages", [False, True])
@pytest.mark.parametrize("name", ["square2d", "sphere3d_fric"])
def test_adapter_steps_the_reference_example(name, stages, tmp_path):
    sc, steps = scenes()[name]
    path = tmp_path / "scene.bin"
    scene_io.write_scene(str(path), sc, steps, steps)
    out = tmp_path / "gpu.bin"
    print(f"adapter_check {name} {steps} steps stages={stages}", flush=True)
    r = subprocess.run([EXE, str(path), str(out)] + (["stages"] if stages else []), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    got = scene_io.read_output(str(out), sc["dim"])[-1]
    ref = scene_io.run_ref(sc, steps)[-1]
    assert np.array_equal(got["valid"], ref["valid"])
    h = 1.0 / np.asarray(sc["N"], float).max()
    for k in ("X", "V", "F", "B", "grid_mass", "grid_velocity"):
        scale = np.abs(ref[k]).max()
        if k == "B":
            scale = max(scale, np.abs(ref["V"]).max() * h)
        err = np.abs(got[k] - ref[k]).max() / scale
        print(name, k, f"{err:.2e}", flush=True)
        assert err <= 1e-9, (k, err)
