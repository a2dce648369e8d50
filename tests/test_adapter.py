"""The drop-in adapter a PhysBAM maintainer adds (physbam_amd/host/MPM_GPU_STEP.h)
compiles and links against the reference's own headers and library
(oracle/ref/Makefile, target adapter_check), and refuses, before it touches
the device, a scene the GPU step does not take: a moving collision object
(MPM_COLLISION_OBJECT::func_frame, Hybrid_Methods/Collisions/
MPM_COLLISION_OBJECT.h:23-24).  CPU only; skipped where the reference
checkout or its built library is absent."""
import os
import subprocess

import pytest

import mpmtest as mt
import scene_io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/Public_Library"
ARCHIVE = os.path.join(ROOT, "oracle", "_ref", "libphysbam_ref.a")
EXE = os.path.join(ROOT, "oracle", "_ref", "adapter_check")

pytestmark = pytest.mark.skipif(not (os.path.isdir(REF) and os.path.isfile(ARCHIVE)),
                                reason="needs the reference checkout and oracle/_ref/libphysbam_ref.a")


def test_adapter_compiles_against_reference():
    r = subprocess.run(["make", "-B", EXE], cwd=os.path.join(ROOT, "oracle", "ref"), capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert os.access(EXE, os.X_OK)


def test_adapter_refuses_moving_collision_objects(tmp_path):
    if not os.access(EXE, os.X_OK):
        pytest.skip("adapter_check not built")
    sq = lambda X: (abs(X[:, 0] - 0.5) <= 0.2) & (abs(X[:, 1] - 0.6) <= 0.2)
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=8, inset=0.05, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(2, 32, prec=8, model=mt.COROTATED, shape=sq, walls=walls, dt=1e-4, seed=3)
    path = tmp_path / "scene.bin"
    scene_io.write_scene(str(path), sc, 3, 3)
    r = subprocess.run([EXE, str(path), str(tmp_path / "out.bin"), "moving"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 3, r.stderr[-1000:]
    assert "moving collision objects" in r.stderr
    assert not (tmp_path / "out.bin").exists()
