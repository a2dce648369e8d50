"""Scene files for the reference harness (oracle/ref/ref_harness.cpp).

Test infrastructure: writes the binary scene layout the harness reads, runs
the harness binary built under oracle/_ref/, and parses its dumps.  The
harness exists only where the reference checkout was available at build time;
callers check ``ref_available()`` first.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def ref_available():
    return os.path.isfile(REF_HARNESS) and os.access(REF_HARNESS, os.X_OK)


def plastic_kind(sc):
    """0 none, 1 MPM_PLASTICITY_CLAMP, 2 MPM_DRUCKER_PRAGER (sc["plasticity"]["kind"])."""
    pl = sc.get("plasticity")
    if not pl:
        return 0
    return 2 if pl.get("kind") == "drucker_prager" else 1


def write_scene(path, sc, nsteps, dump_every, threads=1):
    """Write scene dict ``sc`` (see tests/mpmtest.py make_* helpers)."""
    d = sc["dim"]
    n = int(sc["X"].shape[0])
    N = list(sc["N"]) + [1] * (3 - d)
    lo = list(sc["lo"]) + [0.0] * (3 - d)
    hi = list(sc["hi"]) + [0.0] * (3 - d)
    g = list(sc["gravity"]) + [0.0] * (3 - d)
    walls = sc.get("walls", [])
    pl = sc.get("plasticity")
    with open(path, "wb") as f:
        f.write(b"MPMSCN1\0")
        f.write(struct.pack("<8i", d, sc["prec"], sc["model"], nsteps, dump_every,
                            len(walls), threads, plastic_kind(sc)))
        f.write(struct.pack("<3i", *N))
        f.write(struct.pack("<3d", *lo))
        f.write(struct.pack("<3d", *hi))
        f.write(struct.pack("<3d", *g))
        f.write(struct.pack("<2d", sc["dt"], sc.get("neo_threshold", 0.25)))
        if plastic_kind(sc) == 1:
            f.write(struct.pack("<4d", pl["theta_c"], pl["theta_s"], pl["max_hardening"], pl["hardening_factor"]))
        elif plastic_kind(sc) == 2:
            f.write(struct.pack("<6d", pl["a0"], pl["a1"], pl["a3"], pl["a4"], pl["cohesion"],
                                1.0 if pl.get("volume_correction", True) else 0.0))
        f.write(struct.pack("<q", n))
        for w in walls:
            wl = list(w["lo"]) + [0.0] * (3 - d)
            wh = list(w["hi"]) + [0.0] * (3 - d)
            f.write(struct.pack("<2i", w["type"], 0))
            f.write(struct.pack("<d", w["friction"]))
            f.write(struct.pack("<3d", *wl))
            f.write(struct.pack("<3d", *wh))
        for key in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda"):
            f.write(np.ascontiguousarray(sc[key], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(sc["valid"], dtype=np.uint8).tobytes())


def read_output(path, dim, plastic=0):
    with open(path, "rb") as f:
        buf = f.read()
    assert buf[:7] == b"MPMOUT1"
    ndumps, d, n, nn = struct.unpack_from("<4i", buf, 8)
    assert d == dim
    off = 24
    dumps = []

    def take(count, dtype):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    for _ in range(ndumps):
        (step,) = struct.unpack_from("<i", buf, off)
        off += 4
        r = {"step": step}
        r["X"] = take(n * d, np.float64).reshape(n, d)
        r["V"] = take(n * d, np.float64).reshape(n, d)
        r["F"] = take(n * d * d, np.float64).reshape(n, d * d)
        r["B"] = take(n * d * d, np.float64).reshape(n, d * d)
        r["valid"] = take(n, np.uint8)
        r["grid_mass"] = take(nn, np.float64)
        r["grid_velocity"] = take(nn * d, np.float64).reshape(nn, d)
        r["grid_velocity_save"] = take(nn * d, np.float64).reshape(nn, d)
        r["grid_velocity_friction_save"] = take(nn * d, np.float64).reshape(nn, d)
        if plastic:
            r["Fp"] = take(n * d * d, np.float64).reshape(n, d * d)
            r["mu"] = take(n, np.float64)
            r["lambda"] = take(n, np.float64)
        if plastic == 2:
            for k in ("plastic_def", "rho_F", "cohesion", "vc"):
                r[k] = take(n, np.float64)
        dumps.append(r)
    return dumps


def reflect_args(sc):
    """Command-line tail of the harness for sc["reflection"] (full, friction, sides)."""
    r = sc.get("reflection")
    if not r:
        return []
    return ["reflect", str(1 if r.get("full") else 0), repr(float(r.get("friction", 0.0)))] + \
        [str(int(t)) for t in r["sides"]]


def run_ref(sc, nsteps, dump_every=None, threads=1):
    """Run the reference MPM_DRIVER on scene ``sc``; returns the list of dumps."""
    if dump_every is None:
        dump_every = nsteps
    with tempfile.TemporaryDirectory() as td:
        sp = os.path.join(td, "scene.bin")
        op = os.path.join(td, "out.bin")
        write_scene(sp, sc, nsteps, dump_every, threads)
        subprocess.run([REF_HARNESS, "step", sp, op] + reflect_args(sc), check=True, stdout=subprocess.DEVNULL)
        return read_output(op, sc["dim"], plastic=plastic_kind(sc))


def run_ref_totals(sc, nsteps, threads=1):
    """Conservation totals of the reference over ``nsteps`` steps of scene
    ``sc`` (harness mode "totals"): an array of (3 nsteps + 1) records of 11
    values -- per step the particles after "update-weights", the grid after
    "p2g" and after "forces"; then the particles at the end."""
    with tempfile.TemporaryDirectory() as td:
        sp = os.path.join(td, "scene.bin")
        op = os.path.join(td, "out.bin")
        write_scene(sp, sc, nsteps, 0, threads)
        subprocess.run([REF_HARNESS, "totals", sp, op], check=True, stdout=subprocess.DEVNULL)
        with open(op, "rb") as f:
            (nrec,) = struct.unpack("<i", f.read(4))
            return np.frombuffer(f.read(8 * 11 * nrec), dtype=np.float64).reshape(nrec, 11).copy()


def _limit_args(lim):
    return [str(1 if lim.get("strong_cfl") else 0), str(1 if lim.get("sound_cfl") else 0),
            repr(float(lim.get("cfl_F", 0.5))), repr(float(lim.get("cfl_sound", 0.9)))]


def run_ref_frames(sc, nframes, frame_dt, cfl, min_dt, max_dt, threads=1, lim=None):
    """Run ``nframes`` frames of the reference's adaptive substep loop (with
    Reduce_Dt's strong / sound-speed CFL per ``lim``); returns (final dump, dts)."""
    with tempfile.TemporaryDirectory() as td:
        sp = os.path.join(td, "scene.bin")
        op = os.path.join(td, "out.bin")
        write_scene(sp, sc, 0, 1, threads)
        subprocess.run([REF_HARNESS, "frames", sp, str(nframes), repr(float(frame_dt)), repr(float(cfl)),
                        repr(float(min_dt)), repr(float(max_dt)), op] + (_limit_args(lim) if lim else []),
                       check=True, stdout=subprocess.DEVNULL)
        with open(op, "rb") as f:
            buf = f.read()
    return _frames_output(sc, buf)


def run_ref_limited(sc, nsteps, lim, threads=1):
    """``nsteps`` fixed-dt steps of the reference with Reduce_Dt (lim: strong_cfl,
    sound_cfl, cfl, cfl_F, cfl_sound, min_dt); returns (final dump, dts taken)."""
    with tempfile.TemporaryDirectory() as td:
        sp = os.path.join(td, "scene.bin")
        op = os.path.join(td, "out.bin")
        write_scene(sp, sc, nsteps, nsteps, threads)
        a = _limit_args(lim)
        subprocess.run([REF_HARNESS, "limited", sp, a[0], a[1], repr(float(lim.get("cfl", 1.0))), a[2], a[3],
                        repr(float(lim.get("min_dt", 0.0))), op], check=True, stdout=subprocess.DEVNULL)
        with open(op, "rb") as f:
            buf = f.read()
    return _frames_output(sc, buf)


def run_ref_write(sc, nsteps, threads=1):
    """The reference's particle and restart files after ``nsteps`` steps:
    (deformable_object_particles.gz bytes, restart_data.gz bytes)."""
    with tempfile.TemporaryDirectory() as td:
        sp = os.path.join(td, "scene.bin")
        write_scene(sp, sc, nsteps, nsteps, threads)
        subprocess.run([REF_HARNESS, "write", sp, td], check=True, stdout=subprocess.DEVNULL)
        with open(os.path.join(td, "deformable_object_particles.gz"), "rb") as f:
            parts = f.read()
        with open(os.path.join(td, "restart_data.gz"), "rb") as f:
            restart = f.read()
    return parts, restart


def run_ref_read(prec, dim, particles_gz, restart_gz):
    """The reference's own Read_From_File of a frame directory holding the two
    files: dict X, V, F, B, valid, mass, volume, mu, lambda (Fp, mu0, lambda0
    when stored) and time."""
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "deformable_object_particles.gz"), "wb") as f:
            f.write(particles_gz)
        with open(os.path.join(td, "restart_data.gz"), "wb") as f:
            f.write(restart_gz)
        op = os.path.join(td, "out.bin")
        subprocess.run([REF_HARNESS, "read", str(prec), str(dim), td, op], check=True, stdout=subprocess.DEVNULL)
        with open(op, "rb") as f:
            buf = f.read()
    n, = struct.unpack_from("<i", buf, 0)
    off = 4
    out = {}

    def take(k, count, dtype=np.float64):
        nonlocal off
        a = np.frombuffer(buf, dtype, count=count, offset=off).copy()
        off += a.nbytes
        out[k] = a

    d = dim
    take("X", n * d); take("V", n * d); take("F", n * d * d); take("B", n * d * d)
    take("valid", n, np.uint8)
    take("mass", n); take("volume", n); take("mu", n); take("lambda", n)
    has_fp, = struct.unpack_from("<i", buf, off)
    off += 4
    if has_fp:
        take("Fp", n * d * d); take("mu0", n); take("lambda0", n)
    take("time", 1)
    for k, w in (("X", d), ("V", d), ("F", d * d), ("B", d * d), ("Fp", d * d)):
        if k in out:
            out[k] = out[k].reshape(n, w)
    out["time"] = float(out["time"][0])
    assert off == len(buf)
    return out


def run_ref_poisson(prec, dim, seed, cells, lo, hi, ppc, shapes):
    """The reference's Seed_Particles_Poisson sampling (harness mode
    "poisson"); shapes: list of (type 0 box | 1 sphere, a[3], b[3]).
    Returns a list of point arrays, one per shape."""
    with tempfile.TemporaryDirectory() as td:
        ip = os.path.join(td, "in.bin")
        op = os.path.join(td, "out.bin")
        c = list(cells) + [1] * (3 - len(cells))
        with open(ip, "wb") as f:
            f.write(struct.pack("<4i", prec, dim, seed, len(shapes)))
            f.write(struct.pack("<3i", *c))
            f.write(struct.pack("<d", ppc))
            f.write(struct.pack("<3d", *(list(lo) + [0.0] * (3 - len(lo)))))
            f.write(struct.pack("<3d", *(list(hi) + [1.0] * (3 - len(hi)))))
            for t, a, b in shapes:
                f.write(struct.pack("<i", t))
                f.write(struct.pack("<3d", *(list(a) + [0.0] * (3 - len(a)))))
                f.write(struct.pack("<3d", *(list(b) + [0.0] * (3 - len(b)))))
        subprocess.run([REF_HARNESS, "poisson", ip, op], check=True, stdout=subprocess.DEVNULL)
        with open(op, "rb") as f:
            buf = f.read()
    counts = struct.unpack_from(f"<{len(shapes)}i", buf, 0)
    pts = np.frombuffer(buf, np.float64, offset=4 * len(shapes)).reshape(-1, dim)
    out, o = [], 0
    for c in counts:
        out.append(pts[o:o + c].copy())
        o += c
    return out


def _frames_output(sc, buf):
    assert buf[:7] == b"MPMFRM1"
    nd, d, n, nn = struct.unpack_from("<4i", buf, 8)
    assert d == sc["dim"]
    dts = np.frombuffer(buf, np.float64, count=nd, offset=24).copy()
    rest = b"MPMOUT1\0" + struct.pack("<4i", 1, d, n, nn) + buf[24 + 8 * nd:]
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "dump.bin")
        with open(p, "wb") as f:
            f.write(rest)
        return read_output(p, d)[0], dts


def ref_component(mode, args, payload, outspec):
    """Run a component mode (svd, pstrain, weights) of the harness.

    ``payload`` is the input bytes; ``outspec`` a list of (dtype, count).
    """
    with tempfile.TemporaryDirectory() as td:
        ip = os.path.join(td, "in.bin")
        op = os.path.join(td, "out.bin")
        with open(ip, "wb") as f:
            f.write(payload)
        subprocess.run([REF_HARNESS, mode] + [str(a) for a in args] + [ip, op], check=True)
        with open(op, "rb") as f:
            buf = f.read()
    out, off = [], 0
    for dt, cnt in outspec:
        a = np.frombuffer(buf, dtype=dt, count=cnt, offset=off)
        off += a.nbytes
        out.append(a.copy())
    return out
