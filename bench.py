#!/usr/bin/env python3
"""Headline benchmark: explicit APIC MPM step on one MI355X.

Workload: 3D FP32, 256^3 cells on [0,1]^3, twelve non-overlapping cubes with
sides uniform in [0.08, 0.2] placed in [0.1, 0.9]^3 (box seed 46767, chosen
so the draw holds ~8.0M particles), 8 jittered particles per cell (seed 1234),
fixed-corotated with per-box E log-uniform in [1e3, 1e5] and nu uniform in
[0.2, 0.4], density 1000, gravity -9.8 y, sticky floor below y = 0.05,
dt = 2e-5.  A step is one MPM_DRIVER::Advance_One_Time_Step over all particles
with inputs resident in HBM.

Prints one JSON line (see --help); per-phase GPU times come from HIP events
recorded on the stream the kernels run on, inside the timed region.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from physbam_amd import mpmx  # noqa: E402

METRIC = "particle-steps/s, 8M-particle 3D APIC MPM; fwd+adjoint / fwd time ratio"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E spec peak
BOX_SEED = 46767
SEED = 1234
CELLS = 256
DT = 2e-5

# Algorithmic HBM bytes (FP32, 3D).  Per particle: P2G reads X, V, F, B, m,
# V0, mu, lambda (112 B); G2P reads X, F (48 B) and writes X, V, F, B (96 B).
# Per active grid node: P2G result write 16 B; grid update read 16 + write
# 12 B; G2P read 12 B.
# g2p2g (the fused G2P of one step and P2G of the next): read X, F, m, V0, mu,
# lambda (64 B), write X, V, F, B (96 B); per node read v (12 B), write m, p (16 B)
BYTES_PER_PARTICLE = {"p2g": 112.0, "g2p": 144.0, "g2p2g": 160.0, "grid": 0.0, "step": 256.0}
BYTES_PER_NODE = {"p2g": 16.0, "g2p": 12.0, "g2p2g": 28.0, "grid": 28.0, "step": 56.0}


def headline_boxes():
    lo, hi, E, nu = mpmx.random_boxes(BOX_SEED, nboxes=12, side_min=0.08, side_max=0.2, region_lo=0.1,
                                      region_hi=0.9, E_min=1e3, E_max=1e5, nu_min=0.2, nu_max=0.4, cubes=True)
    return [dict(type=mpmx.SHAPE_BOX, a=lo[i], b=hi[i], density=1000.0, E=E[i], nu=nu[i], v=[0.0, 0.0, 0.0])
            for i in range(len(lo))]


def wall_box(cells, axis, low_side, inset):
    """STANDARD_TESTS_BASE::Add_Walls geometry on [0,1]^3 (Projects/mpm/
    STANDARD_TESTS_BASE.cpp:487-511): the domain thickened by 7 cells, cut at
    `inset` inside the low or high side of `axis`."""
    h = 1.0 / cells
    lo, hi = [-7 * h] * 3, [1 + 7 * h] * 3
    if low_side:
        hi[axis] = inset
    else:
        lo[axis] = 1.0 - inset
    return lo, hi


def floor_wall():
    # the -y side cut at y = 0.05
    return wall_box(CELLS, 1, True, 0.05)


# The 3D FP32 neo-Hookean sphere: 128^3 cells, a sphere of radius 0.25 at
# (0.5, 0.6, 0.5), 8 jittered particles per cell (seed 1234), E 5e4, nu 0.3,
# failure threshold 0.25, density 1000, one velocity for the whole sphere
# drawn uniform in [-0.5, 0.5]^3 (numpy MT19937, seed 1234), gravity, slip
# walls on all six sides inset 3 cells, dt 5e-5, 500 steps.
SPHERE_CELLS = 128
SPHERE_DT = 5e-5
SPHERE_STEPS = 500


def sphere_scene(precision=mpmx.FP32):
    """Particles (dict of arrays, material per particle) and wall boxes of the
    neo-Hookean sphere workload."""
    v = np.random.RandomState(SEED).uniform(-0.5, 0.5, 3)
    sph = [dict(type=mpmx.SHAPE_SPHERE, a=[0.5, 0.6, 0.5], b=[0.25, 0.0, 0.0], density=1000.0, E=5e4, nu=0.3,
                v=list(v))]
    P = mpmx.seed_particles(3, SPHERE_CELLS, [0.0] * 3, [1.0] * 3, sph, ppc_axis=2, seed=SEED, precision=precision)
    inset = 3.0 / SPHERE_CELLS
    walls = [wall_box(SPHERE_CELLS, a, s, inset) for a in range(3) for s in (True, False)]
    return P, walls


def sphere_context(P, walls, precision=mpmx.FP32):
    ctx = mpmx.Context(3, SPHERE_CELLS, [0.0] * 3, [1.0] * 3, precision=precision)
    ctx.set_constitutive(mpmx.NEO_HOOKEAN, 0.25)
    ctx.set_gravity([0.0, -9.8, 0.0])
    for lo, hi in walls:
        ctx.add_collision_box(mpmx.SLIP, 0.0, lo, hi)
    ctx.set_particles(P["X"], P["V"], P["F"], P["B"], P["mass"], P["volume"], P["mu"], P["lam"], P["valid"])
    return ctx


def sphere_throughput(log, steps=SPHERE_STEPS, warmup=10):
    """The neo-Hookean sphere at full size: time of `steps` steps after
    `warmup` untimed ones, inputs resident in HBM."""
    P, walls = sphere_scene()
    n = int(P["X"].shape[0])
    ctx = sphere_context(P, walls)
    del P
    ctx.step_async(SPHERE_DT, warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step_async(SPHERE_DT, steps)
    ctx.synchronize()
    sec = time.perf_counter() - t0
    _, _, nvalid = ctx.particle_stats()
    ctx.close()
    log(f"sphere: {n} particles, {steps} steps in {sec * 1e3:.1f} ms")
    return {"value": n * steps / sec, "unit": "particle-steps/s", "ms_per_step": sec * 1e3 / steps, "particles": n,
            "valid_particles_after": nvalid, "steps": steps,
            "workload": "neo-Hookean sphere r 0.25 at (0.5, 0.6, 0.5), 128^3 grid, 8 ppc, E 5e4, nu 0.3, "
                        "one random velocity in [-0.5, 0.5]^3, gravity, slip walls inset 3 cells, dt 5e-5, FP32"}


def deformed_throughput(log, steps=100, warmup=10):
    """The 12-box scene deformed: per box an affine velocity v_k + A_k (x - c_k)
    (A_k entries U[-20, 20] /s, B = A_k h^2/4) and a pre-stretched F = R
    diag(s) R^T (s U[0.85, 1.2] per axis, R a random rotation), so that the
    stress takes its large-strain branch in most boxes, the small-strain
    series in the others; same grid, material and dt as the headline."""
    shapes = headline_boxes()
    P = mpmx.seed_particles(3, CELLS, [0.0] * 3, [1.0] * 3, shapes, ppc_axis=2, seed=SEED, precision=mpmx.FP32)
    X = P["X"]
    n = int(X.shape[0])
    rng = np.random.default_rng(11)
    h = 1.0 / CELLS
    e = h / 8  # a float32 position may round just outside its box; boxes are at least half a cell apart
    V = np.zeros_like(X)
    F = P["F"].copy()
    B = np.zeros_like(P["B"])
    for sh in shapes:
        m = np.all((X >= np.asarray(sh["a"]) - e) & (X <= np.asarray(sh["b"]) + e), axis=1)
        A = rng.uniform(-20.0, 20.0, (3, 3))
        c = 0.5 * (np.asarray(sh["a"]) + np.asarray(sh["b"]))
        V[m] = (rng.uniform(-0.5, 0.5, 3) + (X[m].astype(np.float64) - c) @ A.T).astype(np.float32)
        B[m] = (A * (h * h / 4)).T.reshape(9).astype(np.float32)  # column-major
        Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        Fk = Q @ np.diag(rng.uniform(0.85, 1.2, 3)) @ Q.T
        F[m] = Fk.T.reshape(9).astype(np.float32)
    P = dict(P, V=V, F=F, B=B)
    ctx = make_context(P)
    del P, X, V, F, B
    ctx.step_async(DT, warmup)
    ctx.synchronize()
    # the launch autotuner on this scene (it restores the particles and time)
    chosen, times = ctx.autotune(DT, 16)
    log(f"deformed: autotuner chose {chosen}")
    t0 = time.perf_counter()
    ctx.step_async(DT, steps)
    ctx.synchronize()
    sec = time.perf_counter() - t0
    _, movers, slow = ctx.bin_stats(with_slow=True)
    _, _, nvalid = ctx.particle_stats()
    ctx.close()
    log(f"deformed: {n} particles, {steps} steps in {sec * 1e3:.1f} ms")
    return {"value": n * steps / sec, "unit": "particle-steps/s", "ms_per_step": sec * 1e3 / steps, "particles": n,
            "valid_particles_after": nvalid, "movers_last_step": movers, "fp64_stress_last_step": slow,
            "steps": steps, "launch_config": chosen, "autotune_ms_per_step": times,
            "workload": "12-box scene, 256^3, per-box affine velocity (A U[-20,20]/s) and pre-stretched F "
                        "(stretches U[0.85,1.2]), fixed-corotated, sticky floor, dt 2e-5, FP32; autotuned launch "
                        "configuration (candidates as for the batched line)"}


def make_context(P):
    ctx = mpmx.Context(3, CELLS, [0.0] * 3, [1.0] * 3, precision=mpmx.FP32)
    ctx.set_constitutive(mpmx.FIXED_COROTATED)
    ctx.set_gravity([0.0, -9.8, 0.0])
    lo, hi = floor_wall()
    ctx.add_collision_box(mpmx.STICK, 0.0, lo, hi)
    ctx.set_particles(P["X"], P["V"], P["F"], P["B"], P["mass"], P["volume"], P["mu"], P["lam"], P["valid"])
    return ctx


HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")
# kernel phase -> rocprof kernel-name prefix of its main kernel
TRAFFIC_KERNELS = {"g2p2g": "void mpmx::g2p2g_kernel<float, 3, false, ", "grid": "void mpmx::grid_kernel<float, 3, false>",
                   "side": "void mpmx::fused_side_kernel<float, 3, false>",
                   "glist": "void mpmx::glist_kernel<float, 3, true>"}


def kernel_sources_digest():
    """SHA-256 over the HIP sources and headers of the library: a traffic
    record measured on other sources does not describe this build."""
    import hashlib

    h = hashlib.sha256()
    cdir = os.path.join(ROOT, "physbam_amd", "csrc")
    for f in sorted(os.listdir(cdir)):
        if f.endswith((".hip", ".h")):
            with open(os.path.join(cdir, f), "rb") as fh:
                h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()


def _pmc_means(path, counter):
    """Mean per dispatch (after each kernel's first) of a rocprofv3 --pmc CSV."""
    import csv
    vals = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                vals.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return {k: (sum(v[1:]) / (len(v) - 1) if len(v) > 1 else v[0]) for k, v in vals.items()}


def write_traffic(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, copy_mib):
    """profiles/traffic.json from four rocprofv3 --pmc runs (FETCH_SIZE and
    WRITE_SIZE of the benchmark, then of --copy-probe MIB): KiB per dispatch
    scaled by what the counters read for a plain copy of known size, with the
    digest of the sources measured."""
    nbytes = copy_mib * (1 << 20)
    cf = max(_pmc_means(copy_fetch_csv, "FETCH_SIZE").items(), key=lambda kv: kv[1])[1]
    cw = max(_pmc_means(copy_write_csv, "WRITE_SIZE").items(), key=lambda kv: kv[1])[1]
    fs, ws = nbytes / (cf * 1024), nbytes / (cw * 1024)
    fk, wk = _pmc_means(fetch_csv, "FETCH_SIZE"), _pmc_means(write_csv, "WRITE_SIZE")
    out = {"sources_sha256": kernel_sources_digest(),
           "copy_probe": {"bytes": nbytes, "fetch_kib": cf, "write_kib": cw, "fetch_scale": fs, "write_scale": ws}}
    for phase, prefix in TRAFFIC_KERNELS.items():
        kf = [k for k in fk if k.startswith(prefix)]
        kw = [k for k in wk if k.startswith(prefix)]
        if kf and kw:
            f, w = fk[kf[0]], wk[kw[0]]
            out[phase] = {"fetch_kib": f, "write_kib": w, "bytes_per_launch": (f * fs + w * ws) * 1024}
    with open(TRAFFIC, "w") as fh:
        json.dump(out, fh, indent=1)
    return out


def cpu_threads():
    """OpenMP threads for the reference CPU step: OMP_NUM_THREADS if set, else
    every core the process sees."""
    return max(1, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))


def reference_time(sc, threads, warm, steps, log=None):
    """Wall time of `steps` MPM_DRIVER::Advance_One_Time_Step calls of the
    reference built under oracle/_ref after `warm` untimed ones (harness mode
    "time": output and statistics off, fixed dt); returns its JSON record."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import scene_io

    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "scene.bin")
        scene_io.write_scene(path, sc, 0, 0, threads)
        r = subprocess.run([HARNESS, "time", path, str(threads), str(warm), str(steps)], capture_output=True,
                           text=True, timeout=1800)
    if r.returncode != 0:
        raise RuntimeError(f"reference harness failed: {r.stderr[-400:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def scene_dict(P, cells, model, walls, dt, gravity=(0.0, -9.8, 0.0), wall_type=mpmx.STICK, dim=3):
    """Scene record of the oracle and the harness from seeded particles."""
    return dict(dim=dim, prec=4 if P["X"].dtype == np.float32 else 8, model=model, N=[cells] * dim,
                lo=[0.0] * dim, hi=[1.0] * dim, gravity=list(gravity)[:dim], dt=dt, neo_threshold=0.25,
                walls=[dict(type=wall_type, friction=0.0, lo=lo, hi=hi) for lo, hi in walls], X=P["X"], V=P["V"],
                F=P["F"], B=P["B"], mass=P["mass"], volume=P["volume"], mu=P["mu"], valid=P["valid"],
                **{"lambda": P["lam"]})


def cpu_baseline(shapes, log):
    """Reference CPU step (the PhysBAM MPM_DRIVER built under oracle/_ref) on
    the whole headline scene: 1 warm-up + 3 timed steps (~10-30 s of CPU work
    at OpenMP threads = OMP_NUM_THREADS, else all cores).  Without the harness
    (a box where the reference was not built), the oracle restatement on one
    cube of the scene, serially."""
    threads = cpu_threads()
    if os.access(HARNESS, os.X_OK):
        P = mpmx.seed_particles(3, CELLS, [0.0] * 3, [1.0] * 3, shapes, ppc_axis=2, seed=SEED,
                                precision=mpmx.FP32)
        n = int(P["X"].shape[0])
        sc = scene_dict(P, CELLS, mpmx.FIXED_COROTATED, [floor_wall()], DT)
        del P
        warm, steps = 1, 3
        log(f"cpu baseline: reference MPM_DRIVER, {n} particles, {threads} threads")
        res = reference_time(sc, threads, warm, steps, log)
        return {"value": res["particle_steps_per_s"], "unit": "particle-steps/s", "cores": threads,
                "nproc": os.cpu_count(), "kind": "reference",
                "sample": f"the whole 12-box scene ({n} particles, 256^3 grid, fixed-corotated, sticky floor, "
                          f"dt 2e-5), {warm} warm-up + {steps} timed MPM_DRIVER::Advance_One_Time_Step calls, "
                          f"FP32, OpenMP threads={threads}"}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import mpmtest

    P = mpmx.seed_particles(3, CELLS, [0.0] * 3, [1.0] * 3, shapes[:1], ppc_axis=2, seed=SEED,
                            precision=mpmx.FP32)
    n = int(P["X"].shape[0])
    sc = scene_dict(P, CELLS, mpmx.FIXED_COROTATED, [floor_wall()], DT)
    log(f"cpu baseline: oracle restatement, {n} particles, 1 thread")
    mpmtest.oracle_run(sc, 1, prec=4, want_grid=False)
    t0 = time.perf_counter()
    mpmtest.oracle_run(sc, 1, prec=4, want_grid=False)
    sec = time.perf_counter() - t0
    return {"value": n / sec, "unit": "particle-steps/s", "cores": 1, "nproc": os.cpu_count(), "kind": "port",
            "sample": f"one cube of the 12-box scene ({n} particles), 1 step of the serial oracle, FP32"}


def cpu_table(log, threads_list):
    """Reference CPU particle-steps/s of every configuration (harness mode
    "time"), for each thread count in threads_list: the 2D square (FP64, 200
    steps), the neo-Hookean sphere (2 warm-up + 20), the 12-box scene (2 + 10),
    the 256 batched scenes (each scene 1 warm-up + 5, summed) and the
    differentiable box's forward step (200)."""
    rows = []

    def row(name, sc, warm, steps, threads):
        r = reference_time(sc, threads, warm, steps)
        log(f"{name}: {r['particles']} particles, {threads} threads: {r['particle_steps_per_s']:.3e} particle-steps/s")
        return r

    for th in threads_list:
        # 2D FP64 square: 64^2, [0.3,0.7]x[0.4,0.8], 4 ppc, fixed corotated E 1e4 nu 0.3, sticky ground y < 0.05
        sq = [dict(type=mpmx.SHAPE_BOX, a=[0.3, 0.4, 0.0], b=[0.7, 0.8, 0.0], density=1000.0, E=1e4, nu=0.3,
                   v=[0.0] * 3)]
        P = mpmx.seed_particles(2, 64, [0.0] * 2, [1.0] * 2, sq, ppc_axis=2, seed=SEED, precision=mpmx.FP64)
        h = 1.0 / 64
        r = row("square2d", scene_dict(P, 64, mpmx.FIXED_COROTATED, [([-7 * h, -7 * h], [1 + 7 * h, 0.05])], 1e-4,
                                       dim=2), 0, 200, th)
        rows.append(dict(r, config="2D square, 64^2, FP64"))
        P, walls = sphere_scene()
        r = row("sphere", scene_dict(P, SPHERE_CELLS, mpmx.NEO_HOOKEAN, walls, SPHERE_DT, wall_type=mpmx.SLIP), 2,
                20, th)
        rows.append(dict(r, config="neo-Hookean sphere, 128^3, FP32"))
        P = mpmx.seed_particles(3, CELLS, [0.0] * 3, [1.0] * 3, headline_boxes(), ppc_axis=2, seed=SEED,
                                precision=mpmx.FP32)
        r = row("12-box", scene_dict(P, CELLS, mpmx.FIXED_COROTATED, [floor_wall()], DT), 2, 10, th)
        rows.append(dict(r, config="12 boxes, 256^3, FP32"))
        del P
        sizes, parts = batched_scenes()
        tot_n, tot_s = 0, 0.0
        for N, P in zip(sizes, parts):
            rr = reference_time(scene_dict(P, N, mpmx.FIXED_COROTATED, [], 1e-4), th, 1, 5)
            tot_n += rr["particles"] * rr["steps"]
            tot_s += rr["seconds"]
        log(f"batched: {len(parts)} scenes, {th} threads: {tot_n / tot_s:.3e} particle-steps/s")
        rows.append(dict(config="256 batched scenes, 32^3..96^3, FP32", threads=th, particles=sum(
            int(P["X"].shape[0]) for P in parts), steps=5, seconds=tot_s, particle_steps_per_s=tot_n / tot_s))
        box = [dict(type=mpmx.SHAPE_BOX, a=[0.35] * 3, b=[0.65] * 3, density=1000.0, E=1e4, nu=0.3, v=[0.0] * 3)]
        P = mpmx.seed_particles(3, 64, [0.0] * 3, [1.0] * 3, box, ppc_axis=2, seed=SEED, precision=mpmx.FP32)
        r = row("diff box", scene_dict(P, 64, mpmx.FIXED_COROTATED, [], 1e-4), 0, 200, th)
        rows.append(dict(r, config="differentiable box forward, 64^3, FP32"))
    return rows


def batched_scenes(nscenes=256, seed=BOX_SEED):
    """The batched workload's scenes: grid N uniform in {32, 48, 64, 96}, one
    sphere of radius U[0.1, 0.3] at the centre, 8 ppc (seed 1234 + scene),
    E log-U[1e3, 1e5], nu U[0.2, 0.4]; returns the grid sizes and the
    per-scene particle dicts."""
    rng = np.random.default_rng(seed)
    sizes, parts = [], []
    for s in range(nscenes):
        N = int(rng.choice([32, 48, 64, 96]))
        r = float(rng.uniform(0.1, 0.3))
        E = float(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
        nu = float(rng.uniform(0.2, 0.4))
        sph = [dict(type=mpmx.SHAPE_SPHERE, a=[0.5, 0.5, 0.5], b=[r, 0.0, 0.0], density=1000.0, E=E, nu=nu,
                    v=[0.0] * 3)]
        sizes.append(N)
        parts.append(mpmx.seed_particles(3, N, [0.0] * 3, [1.0] * 3, sph, ppc_axis=2, seed=SEED + s,
                                         precision=mpmx.FP32))
    return sizes, parts


def batched(log, nscenes=256, steps=100, seed=BOX_SEED):
    """Batched configuration: independent 3D scenes in one context (see
    batched_scenes), fixed corotated, gravity, FP32, dt 1e-4; every step
    advances all scenes with the same launches."""
    t0 = time.time()
    sizes, parts = batched_scenes(nscenes, seed)
    grids = [(N, [0.0] * 3, [1.0] * 3) for N in sizes]
    counts = [int(P["X"].shape[0]) for P in parts]
    n = sum(counts)
    cat = lambda k: np.concatenate([P[k] for P in parts])
    log(f"batched: {nscenes} scenes, {n} particles ({min(counts)}..{max(counts)}), seeded in {time.time() - t0:.1f}s")
    ctx = mpmx.Context.batch(3, grids, precision=mpmx.FP32)
    ctx.set_constitutive(mpmx.FIXED_COROTATED)
    ctx.set_gravity([0.0, -9.8, 0.0])
    arrays = [cat(k) for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lam", "valid")]
    del parts

    def load():
        ctx.set_batch_particles(counts, *arrays)

    dt = 1e-4
    # the launch autotuner picks graph replay and the tile kernels' workgroup
    # count on these scenes (the particles are restored afterwards)
    load()
    chosen, times = ctx.autotune(dt, 16)
    log(f"batched: autotuner chose {chosen} ({', '.join(f'{t:.3f}' for t in times)} ms per step)")

    def timed(graph):
        """steps from the initial state, after 5 untimed ones"""
        ctx.set_launch_config(graph=graph, tile_workgroups=chosen["tile_workgroups"],
                              tile_threads=chosen["tile_threads"])
        load()
        ctx.step(dt, 5)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.step(dt, steps)
        return time.perf_counter() - t1

    sec_off = timed(False)
    sec_on = timed(True)
    sec = sec_on if chosen["graph"] else sec_off
    ctx.close()
    return {"value": n * steps / sec, "unit": "particle-steps/s", "ms_per_step": sec * 1e3 / steps,
            "launch_config": chosen, "autotune_ms_per_step": times,
            "ms_per_step_graph_off": sec_off * 1e3 / steps, "ms_per_step_graph_on": sec_on * 1e3 / steps,
            "scenes": nscenes, "particles": n, "particles_min": min(counts), "particles_max": max(counts),
            "steps": steps,
            "workload": f"{nscenes} independent 3D scenes, N in {{32,48,64,96}}^3, one sphere r U[0.1,0.3], 8 ppc, "
                        "fixed-corotated, E log-U[1e3,1e5], nu U[0.2,0.4], gravity, dt 1e-4, FP32, one context; "
                        "autotuned launch configuration (candidate times in autotune_ms_per_step: graph off, then "
                        "on; tile threads 128, then 256; tile workgroups 0, 1x, 2x, 4x resident)"}


def adjoint_ratio(log, steps=200, checkpoint=20):
    """Differentiable configuration: 64^3 grid, one 0.3^3 box of 8 ppc
    (~56.6k particles), fixed corotated E 1e4, nu 0.3, gravity, dt 1e-4,
    200 steps, checkpoint every 20, FP32; loss |mean x(T) - (0.7, 0.3, 0.5)|^2.
    Returns the time of a taped forward + reverse pass over the time of a
    plain forward pass of the same steps, both after an untimed warm-up of
    each (as in a training loop past its first iteration), each the median of
    three runs."""
    cells, dt = 64, 1e-4
    box = [dict(type=mpmx.SHAPE_BOX, a=[0.35] * 3, b=[0.65] * 3, density=1000.0, E=1e4, nu=0.3, v=[0.0] * 3)]
    P = mpmx.seed_particles(3, cells, [0.0] * 3, [1.0] * 3, box, ppc_axis=2, seed=SEED, precision=mpmx.FP32)
    n = int(P["X"].shape[0])
    ctx = mpmx.Context(3, cells, [0.0] * 3, [1.0] * 3, precision=mpmx.FP32)
    ctx.set_constitutive(mpmx.FIXED_COROTATED)
    ctx.set_gravity([0.0, -9.8, 0.0])
    target = np.array([0.7, 0.3, 0.5])

    def load():
        ctx.set_particles(P["X"], P["V"], P["F"], P["B"], P["mass"], P["volume"], P["mu"], P["lam"], P["valid"])

    load()
    ctx.step(dt, 20)  # warm-up
    # one untimed taped run as a training loop's first iteration: it sizes the
    # tape storage (checkpoints, segment states and grids), kept for later tapes
    load()
    ctx.tape_begin(checkpoint)
    ctx.step(dt, steps)
    ctx.backward(dX=np.zeros((n, 3)))
    def forward():
        load()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.step(dt, steps)
        ctx.synchronize()
        return time.perf_counter() - t0

    def forward_adjoint():
        load()
        ctx.step(dt, 1)
        load()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.tape_begin(checkpoint)
        ctx.step(dt, steps)
        X = ctx.get_particles()["X"].astype(np.float64)
        dX = np.broadcast_to(2 * (X.mean(0) - target) / n, X.shape)
        g = ctx.backward(dX=dX)
        return time.perf_counter() - t0, g

    # the median of three runs of each (the runs last tens of milliseconds)
    t_fwd = float(np.median([forward() for _ in range(3)]))
    runs = [forward_adjoint() for _ in range(3)]
    t_fa = float(np.median([t for t, _ in runs]))
    g = runs[-1][1]
    ctx.close()
    assert np.isfinite(g["V"]).all() and np.isfinite(g["mu"]).all()
    log(f"adjoint: {n} particles, forward {t_fwd * 1e3:.1f} ms, forward+adjoint {t_fa * 1e3:.1f} ms")
    return {"ratio": t_fa / t_fwd, "forward_ms_per_step": t_fwd * 1e3 / steps,
            "forward_adjoint_ms_per_step": t_fa * 1e3 / steps, "particles": n, "steps": steps,
            "checkpoint_every": checkpoint,
            "workload": "64^3 grid, one 0.3^3 box at 8 ppc, fixed-corotated, gravity, dt 1e-4, FP32, "
                        "loss |mean x(T) - (0.7, 0.3, 0.5)|^2, gradients w.r.t. v(0) and Lame parameters"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the reference CPU timing")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint / forward ratio")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-scenes throughput")
    ap.add_argument("--no-sphere", action="store_true", help="skip the neo-Hookean sphere (128^3) throughput")
    ap.add_argument("--no-deformed", action="store_true", help="skip the deformed 12-box scene throughput")
    ap.add_argument("--reorder-interval", type=int, default=16)
    ap.add_argument("--rebin-fraction", type=float, default=1.0 / 128,
                    help="rebuild the bins when more than this fraction of particles left their binned tile")
    ap.add_argument("--rebin-interval", type=int, default=0, help="also rebuild the bins every K steps (0: never)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the forward+adjoint / forward measurement and print its JSON")
    ap.add_argument("--deformed-only", action="store_true",
                    help="only run the deformed 12-box scene (autotuned) and print its JSON")
    ap.add_argument("--cpu-table", type=str, default="", metavar="THREADS",
                    help="only time the reference CPU step of every configuration at these comma-separated "
                         "OpenMP thread counts and print the rows as JSON")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (calibrates HBM byte counters)")
    ap.add_argument("--write-traffic", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="only write profiles/traffic.json from rocprofv3 --pmc counter_collection CSVs of the "
                         "benchmark and of --copy-probe 1024")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)

    if args.write_traffic:
        print(json.dumps(write_traffic(*args.write_traffic, copy_mib=1024)), flush=True)
        return

    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copy probe: 3 copies of {args.copy_probe} MiB")
        return

    if args.cpu_table:
        print(json.dumps(cpu_table(log, [int(t) for t in args.cpu_table.split(",")])), flush=True)
        return

    if args.adjoint_only:
        print(json.dumps(adjoint_ratio(log)), flush=True)
        return

    if args.deformed_only:
        print(json.dumps(deformed_throughput(log)), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    t0 = time.time()
    shapes = headline_boxes()
    P = mpmx.seed_particles(3, CELLS, [0.0] * 3, [1.0] * 3, shapes, ppc_axis=2, seed=SEED, precision=mpmx.FP32)
    n = int(P["X"].shape[0])
    log(f"scene: {len(shapes)} boxes, {n} particles, seeded in {time.time() - t0:.1f}s")

    ctx = make_context(P)
    del P
    ctx.set_reorder_interval(args.reorder_interval)
    ctx.set_rebin_policy(args.rebin_fraction, args.rebin_interval)
    ctx.step_async(DT, args.warmup)
    ctx.synchronize()
    log(f"warm-up: {args.warmup} steps done")

    # the timed region carries HIP events around the fused tile kernel only
    # (timing mode 2); the per-phase breakdown comes from the same number of
    # steps right after it, with every phase's events on
    ctx.enable_timing(2)
    ctx.reset_timing()
    rebins0, _ = ctx.bin_stats()
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.step_async(DT, args.steps)
    ctx.synchronize()
    t2 = time.perf_counter()
    sec = t2 - t1
    log(f"timed: {args.steps} steps in {sec * 1e3:.2f} ms")

    rebins1, movers_last = ctx.bin_stats()
    live_kernel = ctx.phase_time("g2p2g_kernel")
    ctx.enable_timing(1)
    ctx.reset_timing()
    ctx.step_async(DT, args.steps)
    ctx.synchronize()
    phases = {}
    for ph in ("rebin", "bin", "p2g", "grid", "g2p", "g2p2g", "g2p2g_kernel", "reorder", "step"):
        tot, cnt = ctx.phase_time(ph)
        phases[ph] = (tot, cnt)
    ctx.enable_timing(False)
    mass, _ = ctx.get_grid()
    nodes = int(np.count_nonzero(mass))
    _, _, nvalid = ctx.particle_stats()
    del mass
    ctx.close()

    ms_per_step = sec * 1e3 / args.steps
    value = n * args.steps / sec
    # dominant kernel by accumulated time (phase pass)
    kern = max(("p2g", "grid", "g2p", "g2p2g"), key=lambda k: phases[k][0])
    # the g2p2g phase includes the mover side paths: its kernel alone, timed
    # live over the timed region
    ktot, kcnt = live_kernel if kern == "g2p2g" else phases[kern]
    avg_s = ktot / max(kcnt, 1) * 1e-3
    kbytes = BYTES_PER_PARTICLE[kern] * n + BYTES_PER_NODE[kern] * nodes
    achieved = kbytes / avg_s / 1e9
    step_bytes = BYTES_PER_PARTICLE["step"] * n + BYTES_PER_NODE["step"] * nodes
    step_gbs = step_bytes / (sec / args.steps) / 1e9
    # HBM bytes per launch of the dominant kernel from the committed counter
    # runs, only when they were measured on these very sources
    traffic, traffic_note = None, "no counter record"
    if os.path.isfile(TRAFFIC):
        with open(TRAFFIC) as f:
            tr = json.load(f)
        if tr.get("sources_sha256") != kernel_sources_digest():
            traffic_note = "counter record of other sources: not used"
        elif kern in tr:
            traffic, traffic_note = tr[kern]["bytes_per_launch"], "rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated"

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "12-box 3D APIC MPM, 256^3 grid, 8 ppc, fixed-corotated, sticky floor, dt 2e-5",
                   "particles": n, "valid_particles_after": nvalid, "active_nodes": nodes, "grid": [CELLS] * 3,
                   "box_seed": BOX_SEED, "seed": SEED, "parallelism": "single GPU",
                   "reorder_interval": args.reorder_interval, "rebin_fraction": args.rebin_fraction,
                   "rebin_interval": args.rebin_interval, "rebins_timed": rebins1 - rebins0,
                   "movers_last_step": movers_last},
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "bytes_per_launch": kbytes, "avg_launch_ms": avg_s * 1e3},
        "step_roofline": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                          "bytes_per_step": step_bytes},
        "phase_ms_per_step": {k: v[0] / args.steps for k, v in phases.items()},
        "phase_note": "every phase timed with HIP events over the same number of steps right after the timed region "
                      "(those events take stream time of their own)",
        "phase_launches": {k: v[1] for k, v in phases.items()},
        "adjoint_ratio": None,
        "cpu_baseline": None,
    }
    if not args.no_deformed:
        try:
            out["deformed"] = deformed_throughput(log)
        except Exception as e:  # report, never hide the GPU number
            log(f"deformed run failed: {e}")
    if not args.no_sphere:
        try:
            out["sphere"] = sphere_throughput(log)
        except Exception as e:  # report, never hide the GPU number
            log(f"sphere run failed: {e}")
    if not args.no_batched:
        try:
            out["batched"] = batched(log)
        except Exception as e:  # report, never hide the GPU number
            log(f"batched run failed: {e}")
    if not args.no_adjoint:
        try:
            out["adjoint_ratio"] = adjoint_ratio(log)
        except Exception as e:  # report, never hide the GPU number
            log(f"adjoint ratio failed: {e}")
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(shapes, log)
        except Exception as e:  # report, never hide the GPU number
            log(f"cpu baseline failed: {e}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
