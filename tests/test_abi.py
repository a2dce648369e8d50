"""The C ABI library loads and exports every entry point its headers declare
(CPU only: no compute calls), and the host-side scene generator is
deterministic."""
import ctypes
import os
import re

import numpy as np

from physbam_amd import mpmx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(mpmx_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(mpmx.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(mpmx.EXPORTS)


def test_last_error_is_empty_initially():
    assert mpmx.lib().mpmx_last_error() == b""


def test_null_arguments_are_rejected_without_a_device():
    L = mpmx.lib()
    assert L.mpmx_create(None, None) == -1
    assert b"null" in L.mpmx_last_error()
    assert L.mpmx_step(None, ctypes.c_double(1e-3), 1) == -1
    assert L.mpmx_destroy(None) == 0


def test_seeding_is_deterministic_and_counts_match():
    shapes = [dict(type=mpmx.SHAPE_SPHERE, a=[0.5, 0.5, 0.5], b=[0.2, 0, 0], density=1000.0, E=1e4, nu=0.3)]
    a = mpmx.seed_particles(3, 32, [0, 0, 0], [1, 1, 1], shapes, seed=7)
    b = mpmx.seed_particles(3, 32, [0, 0, 0], [1, 1, 1], shapes, seed=7)
    assert np.array_equal(a["X"], b["X"])
    n = a["X"].shape[0]
    # 8 samples per cell inside a sphere of radius 0.2: about 8 * 32^3 * (4/3 pi 0.2^3)
    assert abs(n - 8 * 32 ** 3 * 4 / 3 * np.pi * 0.2 ** 3) < 0.02 * n
    assert (np.linalg.norm(a["X"] - 0.5, axis=1) <= 0.2).all()
    h = 1 / 32
    assert np.allclose(a["volume"], h ** 3 / 8) and np.allclose(a["mass"], 1000 * h ** 3 / 8)
    mu, la = 1e4 / 2.6, 1e4 * 0.3 / (1.3 * 0.4)
    assert np.allclose(a["mu"], mu, rtol=1e-6) and np.allclose(a["lam"], la, rtol=1e-6)
    c = mpmx.seed_particles(3, 32, [0, 0, 0], [1, 1, 1], shapes, seed=8)
    assert not np.array_equal(a["X"][: min(n, len(c["X"]))], c["X"][: min(n, len(c["X"]))])


def test_random_boxes_do_not_overlap():
    lo, hi, E, nu = mpmx.random_boxes(46767)
    assert len(lo) == 12
    assert ((hi - lo) >= 0.08 - 1e-12).all() and ((hi - lo) <= 0.2 + 1e-12).all()
    assert (lo >= 0.1).all() and (hi <= 0.9 + 1e-12).all()
    for i in range(12):
        for j in range(i):
            sep = (hi[i] <= lo[j]) | (hi[j] <= lo[i])
            assert sep.any()
    assert ((E >= 1e3) & (E <= 1e5)).all() and ((nu >= 0.2) & (nu <= 0.4)).all()
