"""Writes the golden vectors under tests/golden/ from the PhysBAM reference.

Runs the reference harness (oracle/_ref/ref_harness, built from the reference
checkout by oracle/ref/Makefile) on small seeded scenes and component inputs,
and stores inputs and outputs as compressed .npz fixtures.  Run in the
authoring container:  python tests/golden/make_golden.py [steps] [components] [frames] [plastic] [totals] ...
(no argument: every group).
"""
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import mpmtest as mt  # noqa: E402
import scene_io  # noqa: E402


def step_cases():
    sq = lambda X: (np.abs(X[:, 0] - 0.5) <= 0.2) & (np.abs(X[:, 1] - 0.6) <= 0.2)
    sph = lambda X: ((X - np.array([0.5, 0.45, 0.5])) ** 2).sum(1) < 0.3 ** 2
    cases = {}
    # The 2D FP64 configuration: 64^2, 0.4x0.4 square at (0.5, 0.6), 4 jittered
    # particles per cell, fixed-corotated E=1e4 nu=0.3, density 1000, v=0,
    # gravity -9.8, sticky ground at y=0.05, dt=1e-4.
    walls = mt.wall_boxes(2, [64, 64], [0, 0], [1, 1], flags=8, inset=0.05, type_=mt.STICK, friction=0.0)
    cases["square2d_f64"] = (mt.make_scene(2, 64, prec=8, model=mt.COROTATED, shape=sq, walls=walls, dt=1e-4,
                                           seed=1234), (1, 20, 200))
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.SLIP, friction=0.0)
    cases["sphere3d_neo_slip_f32"] = (mt.make_scene(3, 16, prec=4, model=mt.NEO_HOOKEAN, shape=sph,
                                                    v0=[0.3, -3.0, 0.2], vrand=1.0, brand=1e-3, frand=0.2,
                                                    walls=walls, dt=1e-3, per_particle_E=True, seed=5), (1, 10))
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.SEPARATE, friction=0.3)
    cases["sphere3d_coro_sep_fric_f64"] = (mt.make_scene(3, 16, prec=8, model=mt.COROTATED, shape=sph,
                                                         v0=[0.3, -3.0, 0.2], vrand=1.0, brand=1e-3, frand=0.2,
                                                         walls=walls, dt=1e-3, per_particle_E=True, seed=6), (1, 10))
    # neo-Hookean with compressed particles (J < 0.25 takes the clamp branch)
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=15, inset=0.1, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(2, 32, prec=8, model=mt.NEO_HOOKEAN, shape=sq, v0=[0.0, -1.0], walls=walls, dt=1e-4, seed=7)
    sc["F"][::3] *= 0.45
    cases["square2d_neo_clamp_f64"] = (sc, (1, 5))
    return cases


def frame_cases():
    """Scenes run with the reference's adaptive substep loop (Compute_Dt, clamp, frame-end rule).

    Each entry: scene, (nframes, frame_dt, cfl, min_dt, max_dt)."""
    sq = lambda X: (np.abs(X[:, 0] - 0.5) <= 0.15) & (np.abs(X[:, 1] - 0.45) <= 0.15)
    sph = lambda X: ((X - np.array([0.5, 0.5, 0.5])) ** 2).sum(1) < 0.25 ** 2
    cases = {}
    # 2D FP64: a square thrown at the sticky floor; the CFL bound (cfl 0.5)
    # sets dt below max_dt, and the frame end forces shortened substeps.
    walls = mt.wall_boxes(2, [64, 64], [0, 0], [1, 1], flags=15, inset=0.05, type_=mt.STICK, friction=0.0)
    cases["throw2d_f64"] = (mt.make_scene(2, 64, prec=8, model=mt.COROTATED, shape=sq, v0=[2.0, -6.0], vrand=0.5,
                                          walls=walls, dt=1e-3, seed=31), (3, 1.0 / 60, 0.5, 1e-6, 1.0 / 60))
    # 3D FP32 neo-Hookean ball against slip walls, max_dt binding at first.
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.SLIP, friction=0.0)
    cases["ball3d_neo_slip_f32"] = (mt.make_scene(3, 16, prec=4, model=mt.NEO_HOOKEAN, shape=sph,
                                                  v0=[0.5, -4.0, 0.3], vrand=0.5, brand=1e-3, frand=0.1,
                                                  walls=walls, dt=1e-3, per_particle_E=True, seed=32),
                                    (2, 1.0 / 48, 0.6, 1e-5, 2e-3))
    return cases


def write_frames():
    out = {}
    for name, (sc, (nf, fdt, cfl, mind, maxd)) in frame_cases().items():
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
            out[f"{name}_in_{k}"] = np.asarray(sc[k])
        meta = {k: sc[k] for k in ("dim", "prec", "model", "N", "lo", "hi", "gravity", "dt", "neo_threshold")}
        meta.update(nframes=nf, frame_dt=fdt, cfl=cfl, min_dt=mind, max_dt=maxd)
        out[f"{name}_meta"] = np.array(json.dumps(meta, default=float))
        out[f"{name}_walls"] = np.array(json.dumps(sc["walls"], default=float))
        r, dts = scene_io.run_ref_frames(sc, nf, fdt, cfl, mind, maxd)
        out[f"{name}_dts"] = dts
        for k in ("X", "V", "F", "B", "valid"):
            out[f"{name}_out_{k}"] = r[k]
        print("frames", name, sc["X"].shape[0], "particles,", len(dts), "substeps, dt range",
              dts.min(), dts.max())
    np.savez_compressed(os.path.join(HERE, "frames.npz"), **out)


def limit_cases():
    """Scenes stepped with Reduce_Dt's strong and sound-speed CFL
    (use_strong_cfl, use_sound_speed_cfl): fixed-dt steps whose dt is lowered
    in some steps and kept in others, one per constitutive model, and one
    frame run whose substeps are lowered inside the frame.

    Each entry: scene, limits, ("steps", n) or ("frames", nframes, frame_dt, cfl, min_dt, max_dt)."""
    sq = lambda X: (np.abs(X[:, 0] - 0.5) <= 0.15) & (np.abs(X[:, 1] - 0.45) <= 0.15)
    sph = lambda X: ((X - 0.5) ** 2).sum(1) < 0.25 ** 2
    cases = {}
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=15, inset=0.1, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(2, 32, prec=8, model=mt.COROTATED, shape=sq, v0=[2.0, -6.0], vrand=3.0, frand=0.05,
                       walls=walls, dt=3.15e-3, seed=51)
    cases["strong2d_f64"] = (sc, dict(strong_cfl=True, cfl=0.02, cfl_F=0.1), ("steps", 6))
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=15, inset=0.1, type_=mt.STICK, friction=0.3)
    sc = mt.make_scene(2, 32, prec=8, model=mt.HENCKY, shape=sq, v0=[1.0, -3.0], vrand=4.0, frand=0.1, walls=walls,
                       dt=1e-3, E=3e5, seed=53)
    cases["both2d_hencky_fric_f64"] = (sc, dict(strong_cfl=True, sound_cfl=True, cfl=0.05, cfl_F=0.05, cfl_sound=0.9,
                                                min_dt=1e-5), ("steps", 6))
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(3, 16, prec=8, model=mt.COROTATED, shape=sph, v0=[0.5, -2.0, 0.3], vrand=0.5, frand=0.1,
                       walls=walls, dt=1e-3, E=1e6, per_particle_E=True, seed=54)
    cases["sound3d_coro_f64"] = (sc, dict(sound_cfl=True, cfl_sound=0.9), ("steps", 3))
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.SLIP, friction=0.0)
    sc = mt.make_scene(3, 16, prec=4, model=mt.NEO_HOOKEAN, shape=sph, v0=[0.5, -2.0, 0.3], vrand=0.5, frand=0.1,
                       walls=walls, dt=2e-3, E=1e6, per_particle_E=True, seed=52)
    cases["sound3d_neo_f32"] = (sc, dict(sound_cfl=True, cfl_sound=0.9), ("steps", 4))
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(3, 16, prec=4, model=mt.HENCKY, shape=sph, v0=[0.5, -2.0, 0.3], vrand=2.0, frand=0.1,
                       walls=walls, dt=1e-3, E=1e5, per_particle_E=True, seed=55)
    cases["strong3d_hencky_f32"] = (sc, dict(strong_cfl=True, sound_cfl=True, cfl=0.002, cfl_F=0.05, cfl_sound=0.9),
                                    ("steps", 4))
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=15, inset=0.1, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(2, 32, prec=8, model=mt.COROTATED, shape=sq, v0=[1.0, -4.0], vrand=1.0, walls=walls, dt=1e-3,
                       E=2e5, seed=56)
    cases["frames2d_f64"] = (sc, dict(strong_cfl=True, sound_cfl=True, cfl_F=0.1, cfl_sound=0.9),
                             ("frames", 2, 1.0 / 240, 0.5, 1e-6, 1.0 / 240))
    return cases


def write_limits():
    out = {}
    for name, (sc, lim, run) in limit_cases().items():
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
            out[f"{name}_in_{k}"] = np.asarray(sc[k])
        meta = {k: sc[k] for k in ("dim", "prec", "model", "N", "lo", "hi", "gravity", "dt", "neo_threshold")}
        meta.update(limits=lim, run=list(run))
        out[f"{name}_meta"] = np.array(json.dumps(meta, default=float))
        out[f"{name}_walls"] = np.array(json.dumps(sc["walls"], default=float))
        if run[0] == "steps":
            r, dts = scene_io.run_ref_limited(sc, run[1], lim)
        else:
            r, dts = scene_io.run_ref_frames(sc, *run[1:], lim=lim)
        out[f"{name}_dts"] = dts
        for k in ("X", "V", "F", "B", "valid"):
            out[f"{name}_out_{k}"] = r[k]
        print("limits", name, sc["X"].shape[0], "particles, dt / requested", np.round(dts / sc["dt"], 4))
    np.savez_compressed(os.path.join(HERE, "limits.npz"), **out)


POISSON_CASES = {
    # name: (prec, dim, cells, particles per cell, shapes (type 0 box | 1 sphere, a, b))
    "poisson2d_f64": (8, 2, [32, 32], 4.0, [(1, [0.5, 0.5, 0.0], [0.2, 0.0, 0.0]),
                                            (0, [0.1, 0.1, 0.0], [0.3, 0.25, 0.0])]),
    "poisson2d_f32_ppc6": (4, 2, [40, 24], 6.0, [(1, [0.4, 0.6, 0.0], [0.25, 0.0, 0.0])]),
    "poisson3d_f32": (4, 3, [16, 16, 16], 8.0, [(1, [0.5, 0.5, 0.5], [0.25, 0.0, 0.0]),
                                                (0, [0.1, 0.1, 0.1], [0.3, 0.2, 0.35])]),
    "poisson3d_f64": (8, 3, [16, 12, 16], 8.0, [(0, [0.2, 0.3, 0.1], [0.6, 0.5, 0.4])]),
}


def write_poisson():
    out = {}
    for name, (prec, dim, cells, ppc, shapes) in POISSON_CASES.items():
        pts = scene_io.run_ref_poisson(prec, dim, 1234, cells, [0.0] * dim, [1.0] * dim, ppc, shapes)
        out[f"{name}_meta"] = np.array(json.dumps(dict(prec=prec, dim=dim, cells=cells, ppc=ppc, seed=1234,
                                                       shapes=[[t, list(a), list(b)] for t, a, b in shapes])))
        for i, p in enumerate(pts):
            out[f"{name}_shape{i}_X"] = p
        print("poisson", name, [len(p) for p in pts], "points")
    np.savez_compressed(os.path.join(HERE, "poisson.npz"), **out)


def reflect_cases():
    """Scenes with reflection boundary conditions (-use_reflect, with
    -use_reflect_friction for the full variant) instead of wall objects:
    particles fill the domain up to the sides, fast enough that some leave it
    and are reflected back, one side left free.

    Each entry: scene, steps."""
    fill = lambda X: np.all((X > 0.003) & (X < 0.997), axis=1)
    shell = lambda X: fill(X) & ~np.all((X > 0.2) & (X < 0.8), axis=1)
    cases = {}
    sc = mt.make_scene(2, 16, prec=8, model=mt.COROTATED, shape=fill, v0=[-1.0, -2.0], vrand=8.0, frand=0.05, dt=1e-3,
                       seed=71)
    sc["reflection"] = dict(full=False, friction=0.0, sides=[mt.SLIP_BC, mt.NOSLIP_BC, mt.NOSLIP_BC, mt.FREE])
    cases["reflect2d_f64"] = (sc, 6)
    sc = mt.make_scene(2, 16, prec=8, model=mt.NEO_HOOKEAN, shape=fill, v0=[1.0, -2.0], vrand=8.0, frand=0.05,
                       dt=1e-3, seed=72)
    sc["reflection"] = dict(full=True, friction=0.4,
                            sides=[mt.SEPARATE_BC, mt.SLIP_BC, mt.SEPARATE_BC, mt.NOSLIP_BC])
    cases["reflect2d_full_f64"] = (sc, 6)
    sc = mt.make_scene(3, 8, prec=4, model=mt.COROTATED, shape=shell, v0=[0.5, -2.0, 1.0], vrand=8.0, frand=0.05,
                       dt=1e-3, seed=73, per_particle_E=True)
    sc["reflection"] = dict(full=False, friction=0.0, sides=[mt.SLIP_BC, mt.NOSLIP_BC, mt.SEPARATE_BC, mt.FREE,
                                                             mt.NOSLIP_BC, mt.SLIP_BC])
    cases["reflect3d_f32"] = (sc, 4)
    sc = mt.make_scene(3, 8, prec=8, model=mt.COROTATED, shape=shell, v0=[0.5, -2.0, 1.0], vrand=8.0, frand=0.05,
                       dt=1e-3, seed=74)
    sc["reflection"] = dict(full=True, friction=0.3, sides=[mt.SEPARATE_BC] * 4 + [mt.NOSLIP_BC, mt.SLIP_BC])
    cases["reflect3d_full_f64"] = (sc, 4)
    return cases


def write_reflect():
    out = {}
    for name, (sc, nsteps) in reflect_cases().items():
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
            out[f"{name}_in_{k}"] = np.asarray(sc[k])
        meta = {k: sc[k] for k in ("dim", "prec", "model", "N", "lo", "hi", "gravity", "dt", "neo_threshold")}
        meta.update(nsteps=nsteps, reflection=sc["reflection"])
        out[f"{name}_meta"] = np.array(json.dumps(meta, default=float))
        out[f"{name}_walls"] = np.array(json.dumps(sc["walls"], default=float))
        for n in (1, nsteps):
            r = scene_io.run_ref(sc, n)[-1]
            for k in ("X", "V", "F", "B", "valid"):
                out[f"{name}_s{n}_{k}"] = r[k]
        free = dict(sc, reflection=dict(sc["reflection"], sides=[0] * (2 * sc["dim"])))
        kept = int(r["valid"].sum()) - int(scene_io.run_ref(free, nsteps)[-1]["valid"].sum())
        print("reflect", name, sc["X"].shape[0], "particles,", kept, "kept by reflection")
    np.savez_compressed(os.path.join(HERE, "reflect.npz"), **out)


def io_cases():
    """Scenes whose particle and restart files the reference writes after a
    few steps (MPM_EXAMPLE::Write_Output_Files' particle part).

    Each entry: scene, steps."""
    sq = lambda X: (np.abs(X[:, 0] - 0.5) <= 0.15) & (np.abs(X[:, 1] - 0.45) <= 0.15)
    sph = lambda X: ((X - 0.5) ** 2).sum(1) < 0.25 ** 2
    cases = {}
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=15, inset=0.1, type_=mt.STICK, friction=0.0)
    cases["io2d_f64"] = (mt.make_scene(2, 32, prec=8, model=mt.COROTATED, shape=sq, v0=[1.0, -3.0], vrand=0.5,
                                       frand=0.05, walls=walls, dt=5e-4, seed=61), 3)
    walls = mt.wall_boxes(3, [12] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.SLIP, friction=0.0)
    cases["io3d_f32"] = (mt.make_scene(3, 12, prec=4, model=mt.NEO_HOOKEAN, shape=sph, v0=[0.3, -2.0, 0.1],
                                       vrand=0.5, frand=0.05, walls=walls, dt=5e-4, per_particle_E=True, seed=62), 2)
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=15, inset=0.1, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(2, 32, prec=8, model=mt.COROTATED, shape=sq, v0=[0.5, -3.0], vrand=0.5, frand=0.08,
                       walls=walls, dt=2e-4, seed=63)
    sc["plasticity"] = dict(theta_c=0.025, theta_s=0.0075, max_hardening=5.0, hardening_factor=10.0)
    cases["io2d_snow_f64"] = (sc, 2)
    return cases


def write_io():
    out = {}
    for name, (sc, nsteps) in io_cases().items():
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
            out[f"{name}_in_{k}"] = np.asarray(sc[k])
        meta = {k: sc[k] for k in ("dim", "prec", "model", "N", "lo", "hi", "gravity", "dt", "neo_threshold")}
        meta.update(nsteps=nsteps)
        if "plasticity" in sc:
            meta["plasticity"] = sc["plasticity"]
        out[f"{name}_meta"] = np.array(json.dumps(meta, default=float))
        out[f"{name}_walls"] = np.array(json.dumps(sc["walls"], default=float))
        parts, restart = scene_io.run_ref_write(sc, nsteps)
        out[f"{name}_particles_gz"] = np.frombuffer(parts, np.uint8)
        out[f"{name}_restart_gz"] = np.frombuffer(restart, np.uint8)
        r = scene_io.run_ref_read(sc["prec"], sc["dim"], parts, restart)
        for k, v in r.items():
            out[f"{name}_read_{k}"] = np.asarray(v)
        print("io", name, sc["X"].shape[0], "particles,", len(parts), "bytes gz, time", r["time"])
    np.savez_compressed(os.path.join(HERE, "io.npz"), **out)


def plastic_cases():
    """Scenes with MPM_PLASTICITY_CLAMP (Add_Clamped_Plasticity); F is perturbed
    so that singular values leave [1 - theta_c, 1 + theta_s] from the first step."""
    sq = lambda X: (np.abs(X[:, 0] - 0.5) <= 0.2) & (np.abs(X[:, 1] - 0.45) <= 0.2)
    sph = lambda X: ((X - np.array([0.5, 0.45, 0.5])) ** 2).sum(1) < 0.3 ** 2
    cases = {}
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=15, inset=0.1, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(2, 32, prec=8, model=mt.COROTATED, shape=sq, v0=[0.5, -3.0], vrand=0.5, frand=0.08,
                       walls=walls, dt=2e-4, seed=41)
    sc["plasticity"] = dict(theta_c=0.025, theta_s=0.0075, max_hardening=5.0, hardening_factor=10.0)
    cases["snow2d_f64"] = (sc, (1, 10))
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.SLIP, friction=0.0)
    sc = mt.make_scene(3, 16, prec=4, model=mt.COROTATED, shape=sph, v0=[0.3, -3.0, 0.2], vrand=1.0, brand=1e-3,
                       frand=0.08, walls=walls, dt=5e-4, per_particle_E=True, seed=42)
    sc["plasticity"] = dict(theta_c=0.02, theta_s=0.01, max_hardening=3.0, hardening_factor=10.0)
    cases["snow3d_f32"] = (sc, (1, 8))
    # sand: Drucker-Prager on the Hencky-strain model (Add_Drucker_Prager), a
    # block dropped on a sticky floor; friction angle hardening a0..a4 of
    # Klar et al. (the reference's default case constants) and no cohesion
    walls = mt.wall_boxes(2, [32, 32], [0, 0], [1, 1], flags=15, inset=0.1, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(2, 32, prec=8, model=mt.HENCKY, shape=sq, v0=[0.5, -3.0], vrand=1.0, frand=0.06,
                       walls=walls, dt=2e-4, seed=43)
    sc["plasticity"] = dict(kind="drucker_prager", a0=35.0, a1=0.0, a3=0.2, a4=10.0, cohesion=0.0,
                            volume_correction=True)
    cases["sand2d_f64"] = (sc, (1, 10))
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=63, inset=0.12, type_=mt.STICK, friction=0.0)
    sc = mt.make_scene(3, 16, prec=4, model=mt.HENCKY, shape=sph, v0=[0.3, -3.0, 0.2], vrand=1.0, frand=0.06,
                       walls=walls, dt=5e-4, per_particle_E=True, seed=44)
    sc["plasticity"] = dict(kind="drucker_prager", a0=30.0, a1=9.0, a3=0.2, a4=10.0, cohesion=100.0,
                            volume_correction=True)
    cases["sand3d_f32"] = (sc, (1, 8))
    return cases


def write_plastic():
    out = {}
    for name, (sc, dumps) in plastic_cases().items():
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
            out[f"{name}_in_{k}"] = np.asarray(sc[k])
        meta = {k: sc[k] for k in ("dim", "prec", "model", "N", "lo", "hi", "gravity", "dt", "neo_threshold",
                                   "plasticity")}
        extra = ("plastic_def", "rho_F", "cohesion", "vc") if scene_io.plastic_kind(sc) == 2 else ()
        out[f"{name}_meta"] = np.array(json.dumps(meta, default=float))
        out[f"{name}_walls"] = np.array(json.dumps(sc["walls"], default=float))
        for n in dumps:
            r = scene_io.run_ref(sc, n)[-1]
            for k in ("X", "V", "F", "B", "valid", "Fp", "mu", "lambda") + extra:
                out[f"{name}_s{n}_{k}"] = r[k]
            moved = (np.abs(r["Fp"] - np.tile(np.eye(sc["dim"]).reshape(-1), (len(r["Fp"]), 1))).max(1) > 0).sum()
            print("plastic", name, "steps", n, "particles with Fp != I:", moved, "of", len(r["Fp"]))
    np.savez_compressed(os.path.join(HERE, "plastic.npz"), **out)


def totals_cases():
    """Scenes for the conservation totals (angular momentum and energies).

    Each entry: scene, steps."""
    cases = {}
    # the reference's rotating sphere (FP64, fixed corotated) and rotating
    # circle (neo-Hookean), no gravity: angular momentum is conserved
    cases["rot3d_coro_f64"] = (mt.make_rotating(3, 16, prec=8, model=mt.COROTATED, omega=[0.4, 0.3, -0.2],
                                                dt=2e-3, seed=41), 6)
    cases["rot2d_neo_f64"] = (mt.make_rotating(2, 32, prec=8, model=mt.NEO_HOOKEAN, omega=[3.0], dt=1e-3,
                                               seed=42), 6)
    # FP32 with gravity, a sticky floor, perturbed F and the Hencky model on
    # an FP64 scene: the potential energy terms of every model
    walls = mt.wall_boxes(3, [16] * 3, [0] * 3, [1] * 3, flags=4, inset=0.12, type_=mt.STICK, friction=0.0)
    sc = mt.make_rotating(3, 16, prec=4, model=mt.NEO_HOOKEAN, omega=[0.0, 2.0, 1.0], dt=1e-3, seed=43,
                          gravity=[0.0, -9.8, 0.0], walls=walls)
    sc["F"] = (sc["F"] + 0.1 * (np.random.default_rng(43).random(sc["F"].shape) - 0.5)).astype(np.float32).astype(
        np.float64)
    cases["rot3d_neo_floor_f32"] = (sc, 4)
    sc = mt.make_rotating(2, 32, prec=8, model=2, omega=[1.0], dt=1e-3, seed=44, gravity=[0.0, -9.8])
    sc["F"] = sc["F"] + 0.05 * (np.random.default_rng(44).random(sc["F"].shape) - 0.5)
    cases["rot2d_hencky_f64"] = (sc, 3)
    return cases


def write_totals():
    out = {}
    for name, (sc, steps) in totals_cases().items():
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
            out[f"{name}_in_{k}"] = np.asarray(sc[k])
        meta = {k: sc[k] for k in ("dim", "prec", "model", "N", "lo", "hi", "gravity", "dt", "neo_threshold")}
        meta["steps"] = steps
        out[f"{name}_meta"] = np.array(json.dumps(meta, default=float))
        out[f"{name}_walls"] = np.array(json.dumps(sc["walls"], default=float))
        out[f"{name}_totals"] = scene_io.run_ref_totals(sc, steps)
        print("totals", name, sc["X"].shape[0], "particles,", steps, "steps")
    np.savez_compressed(os.path.join(HERE, "totals.npz"), **out)


def main():
    if not scene_io.ref_available():
        sys.exit("reference harness missing: build it with oracle/ref/Makefile")
    groups = set(sys.argv[1:]) or {"steps", "components", "frames", "plastic", "limits", "io", "reflect", "poisson",
                                   "totals"}
    if "totals" in groups:
        write_totals()
    if "frames" in groups:
        write_frames()
    if "plastic" in groups:
        write_plastic()
    if "limits" in groups:
        write_limits()
    if "io" in groups:
        write_io()
    if "reflect" in groups:
        write_reflect()
    if "poisson" in groups:
        write_poisson()
    if "steps" in groups:
        write_steps()
    if "components" in groups:
        write_components()


def write_steps():
    for name, (sc, dumps) in step_cases().items():
        out = {}
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda", "valid"):
            out["in_" + k] = np.asarray(sc[k])
        meta = {k: sc[k] for k in ("dim", "prec", "model", "N", "lo", "hi", "gravity", "dt", "neo_threshold")}
        out["meta"] = np.array(json.dumps(meta, default=float))
        out["walls"] = np.array(json.dumps(sc["walls"], default=float))
        for j, n in enumerate(dumps):
            r = scene_io.run_ref(sc, n)[-1]
            keys = ["X", "V", "F", "B", "valid"]
            if j == 0:  # grid state after the first step only, to keep the fixture small
                keys += ["grid_mass", "grid_velocity", "grid_velocity_save", "grid_velocity_friction_save"]
            for k in keys:
                out[f"s{n}_{k}"] = r[k]
        np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), **out)
        print("wrote", name, sc["X"].shape[0], "particles, dumps", dumps)


def write_components():
    rng = np.random.default_rng(2024)
    comp = {}
    for dim in (2, 3):
        n = 512
        A = np.tile(np.eye(dim).reshape(-1), (n, 1)) + 0.6 * (rng.random((n, dim * dim)) - 0.5)
        A[n // 2:] = rng.normal(size=(n // 2, dim * dim)) * rng.choice([1e-3, 1.0, 30.0], size=(n // 2, 1))
        special = [np.eye(dim), -np.eye(dim), np.diag([1.0] * (dim - 1) + [-1.0]), 2 * np.eye(dim),
                   np.diag([3.0] + [1.0] * (dim - 2) + [1e-6]), np.zeros((dim, dim))]
        sp = np.array([m.T.reshape(-1) for m in special])
        A = np.vstack([sp, A])
        for prec, dt in ((8, np.float64), (4, np.float32)):
            Ain = A.astype(dt).astype(np.float64)
            U, S, V = scene_io.ref_component(
                "svd", [prec, dim], struct.pack("<q", len(Ain)) + Ain.tobytes(),
                [(np.float64, Ain.size), (np.float64, len(Ain) * dim), (np.float64, Ain.size)])
            comp[f"svd{dim}_p{prec}_A"] = Ain
            comp[f"svd{dim}_p{prec}_U"] = U.reshape(-1, dim * dim)
            comp[f"svd{dim}_p{prec}_S"] = S.reshape(-1, dim)
            comp[f"svd{dim}_p{prec}_V"] = V.reshape(-1, dim * dim)
        # P_hat(sigma): sorted singular values, some inverted, some compressed
        m = 400
        S = np.sort(rng.uniform(0.05, 1.8, (m, dim)), axis=1)[:, ::-1].copy()
        S[::7, -1] *= -1
        S[::5] *= 0.4
        mu, la = mt.lame(rng.uniform(1e3, 1e5, m), rng.uniform(0.2, 0.4, m))
        for model in (0, 1):
            for prec, dt in ((8, np.float64), (4, np.float32)):
                Sin, mui, lai = (x.astype(dt).astype(np.float64) for x in (S, mu, la))
                (P,) = scene_io.ref_component("pstrain", [prec, dim, model, 0.25],
                                              struct.pack("<q", m) + Sin.tobytes() + mui.tobytes() + lai.tobytes(),
                                              [(np.float64, m * dim)])
                comp[f"pstrain{dim}_m{model}_p{prec}_S"] = Sin
                comp[f"pstrain{dim}_m{model}_p{prec}_mu"] = mui
                comp[f"pstrain{dim}_m{model}_p{prec}_la"] = lai
                comp[f"pstrain{dim}_m{model}_p{prec}_P"] = P.reshape(m, dim)
        X = rng.uniform(0.0, 1.0, (300, dim))
        for prec, dt in ((8, np.float64), (4, np.float32)):
            Xin = X.astype(dt).astype(np.float64)
            base, w, dw = scene_io.ref_component(
                "weights", [prec, dim, 16, 0.0, 1.0], struct.pack("<q", len(Xin)) + Xin.tobytes(),
                [(np.int32, len(Xin) * dim), (np.float64, len(Xin) * 3 * dim), (np.float64, len(Xin) * 3 * dim)])
            comp[f"weights{dim}_p{prec}_X"] = Xin
            comp[f"weights{dim}_p{prec}_base"] = base.reshape(-1, dim)
            comp[f"weights{dim}_p{prec}_w"] = w.reshape(-1, 3, dim)
            comp[f"weights{dim}_p{prec}_dw"] = dw.reshape(-1, 3, dim)
    np.savez_compressed(os.path.join(HERE, "components.npz"), **comp)
    print("wrote components")


if __name__ == "__main__":
    main()
