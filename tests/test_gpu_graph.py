"""Fused steps replayed from a captured hipGraph (mpmx_set_launch_config with
graph on) and the launch autotuner (mpmx_autotune).

A graph replays four fused steps at a time.  Each test runs the same steps
eagerly in another context and compares after every replay, with the
program's own eager work (particle downloads, stream synchronisation)
between replays: FP64 to 1e-12 and FP32 to 1e-5 relative (max|a - b| /
max|b| for X, V, F; B by the velocity its difference carries into the next
P2G; the grid sums' float atomics make the order of additions differ between
any two runs).  Also with bins rebuilt every few
steps (each rebuild re-captures the graph), with persistent tile workgroups
and both workgroup sizes of the fused kernel,
with wall friction, and in a batch of scenes; the autotuner leaves the
particles and the time as they were and picks one of its candidates.
"""
import numpy as np
import pytest

import mpmtest as mt
from physbam_amd import mpmx
from test_gpu_step import gpu_context, rel

pytestmark = pytest.mark.gpu

sph = lambda X: ((X - np.array([0.5, 0.55, 0.5])) ** 2).sum(1) < 0.3 ** 2


def scene(prec, friction=0.0, N=24):
    walls = mt.wall_boxes(3, [N] * 3, [0] * 3, [1] * 3, flags=63, inset=0.15,
                          type_=mt.SLIP if friction else mt.STICK, friction=friction)
    return mt.make_scene(3, N, prec=prec, model=mt.COROTATED, shape=sph, v0=[0.3, -1.0, 0.2], vrand=1.0,
                         brand=1e-3, frand=0.05, walls=walls, dt=1e-3, per_particle_E=True)


def check(a, b, tol, tag, dx=1.0 / 32):
    """X, V, F relative to their largest magnitude; B by the velocity its
    difference carries into the next P2G (relative to the largest speed; dx
    the finest cell size present)."""
    for k in ("X", "V", "F"):
        e = rel(a[k], b[k])
        assert e <= tol, (tag, k, e)
    e = mt.b_velocity_error(a["B"], b["B"], b["V"], dx)
    assert e <= tol, (tag, "B", e)
    assert np.array_equal(a["valid"], b["valid"]), tag


@pytest.mark.parametrize("prec,friction,tile_wg,rebin_every,threads", [
    (8, 0.0, 0, 0, 0),
    (4, 0.0, 0, 0, 256),
    (8, 0.3, 0, 0, 0),
    (8, 0.0, 0, 5, 256),
    (4, 0.0, 512, 3, 128),
])
def test_graph_replays_match_eager_steps(prec, friction, tile_wg, rebin_every, threads):
    sc = scene(prec, friction)
    tol = 1e-12 if prec == 8 else 1e-5
    eager, graph = gpu_context(sc), gpu_context(sc)
    try:
        for c in (eager, graph):
            c.set_rebin_policy(1.0 / 128, rebin_every)
        graph.set_launch_config(graph=True, tile_workgroups=tile_wg, tile_threads=threads)
        assert graph.launch_config() == {"graph": True, "tile_workgroups": tile_wg, "tile_threads": threads or 128}
        # the first four steps run one by one (the pipeline starts, the step
        # count reaches a multiple of four); every later call is whole groups
        for k, n in enumerate((4, 4, 8, 4, 8, 12)):
            eager.step(sc["dt"], n)
            graph.step(sc["dt"], n)
            a, b = eager.get_particles(), graph.get_particles()
            check(b, a, tol, f"call {k}")
        assert graph.time == pytest.approx(eager.time, rel=1e-15)
        reps = graph.phase_time("graph")[1]
        print(f"prec {prec} friction {friction} wg {tile_wg} rebin {rebin_every}: {reps} graph replays", flush=True)
        assert reps == 9  # 36 of the 40 steps ran from the graph
        assert eager.phase_time("graph")[1] == 0
    finally:
        eager.close()
        graph.close()


def test_graph_replays_in_a_batch():
    """Four scenes of other sizes in one batch context: the graph path against
    the eager path, scene by scene."""
    rng = np.random.default_rng(9)
    sizes = [16, 24, 20, 32]
    parts = []
    for i, N in enumerate(sizes):
        s = mt.make_scene(3, N, prec=4, model=mt.COROTATED, shape=sph, v0=list(rng.uniform(-1, 1, 3)), vrand=0.5,
                          walls=[], dt=1e-3, per_particle_E=True, seed=100 + i)
        parts.append(s)
    counts = [p["X"].shape[0] for p in parts]
    cat = lambda k: np.concatenate([p[k] for p in parts])
    ctxs = []
    for use_graph in (False, True):
        c = mpmx.Context.batch(3, [(N, [0.0] * 3, [1.0] * 3) for N in sizes], precision=mpmx.FP32)
        c.set_constitutive(mpmx.FIXED_COROTATED)
        c.set_gravity([0.0, -9.8, 0.0])
        c.set_batch_particles(counts, cat("X"), cat("V"), cat("F"), cat("B"), cat("mass"), cat("volume"), cat("mu"),
                              cat("lambda"), cat("valid"))
        c.set_launch_config(graph=use_graph)
        ctxs.append(c)
    try:
        for n in (4, 8, 8):
            for c in ctxs:
                c.step(1e-3, n)
            a, b = ctxs[0].get_particles(), ctxs[1].get_particles()
            check(b, a, 1e-5, f"batch after {n}")
        assert ctxs[1].phase_time("graph")[1] >= 4
    finally:
        for c in ctxs:
            c.close()


def test_autotune_restores_the_state_and_picks_a_candidate():
    sc = scene(4)
    ctx, ref = gpu_context(sc), gpu_context(sc)
    try:
        ctx.step(sc["dt"], 3)
        ref.step(sc["dt"], 3)
        before, t = ctx.get_particles(), ctx.time
        chosen, times = ctx.autotune(sc["dt"], 8)
        print("autotune:", chosen, [f"{x:.3f}" for x in times], flush=True)
        assert len(times) == 16 and all(x > 0 for x in times)
        assert ctx.launch_config() == chosen
        assert chosen["tile_workgroups"] >= 0
        after = ctx.get_particles()
        for k in ("X", "V", "F", "B", "valid"):
            assert np.array_equal(after[k], before[k]), k
        assert ctx.time == t
        # stepping on with the chosen configuration matches a context that never tuned
        ctx.step(sc["dt"], 8)
        ref.step(sc["dt"], 8)
        check(ctx.get_particles(), ref.get_particles(), 1e-5, "after autotune")
    finally:
        ctx.close()
        ref.close()


def test_launch_config_is_checked():
    sc = scene(4, N=16)
    ctx = gpu_context(sc)
    try:
        with pytest.raises(mpmx.MpmxError):
            ctx.set_launch_config(graph=True, tile_workgroups=-1)
        with pytest.raises(mpmx.MpmxError):
            ctx.set_launch_config(tile_threads=192)
        with pytest.raises(mpmx.MpmxError) as e:
            ctx.autotune(sc["dt"], 4)  # fewer than 8 trial steps
        assert e.value.code == -1
        ctx.enable_grid_stats()
        with pytest.raises(mpmx.MpmxError) as e:
            ctx.autotune(sc["dt"], 8)  # grid stats run the unfused step
        assert e.value.code == -3
    finally:
        ctx.close()


def test_kernel_only_timing():
    """Timing mode 2 records the fused tile kernel alone (the phases around it
    stay untimed) and keeps the steps off the graph, whose replays record no
    events; other modes are rejected."""
    sc = scene(4, N=16)
    ctx = gpu_context(sc)
    try:
        ctx.set_launch_config(graph=True)
        ctx.enable_timing(2)
        ctx.reset_timing()
        ctx.step(sc["dt"], 6)
        ms, cnt = ctx.phase_time("g2p2g_kernel")
        assert cnt == 6 and ms > 0
        for ph in ("grid", "bin", "g2p2g", "step"):
            assert ctx.phase_time(ph)[1] == 0, ph
        assert ctx.phase_time("graph")[1] == 0
        ctx.enable_timing(False)
        ctx.reset_timing()
        ctx.step(sc["dt"], 8)
        assert ctx.phase_time("g2p2g_kernel")[1] == 0
        assert ctx.phase_time("graph")[1] >= 1
        with pytest.raises(mpmx.MpmxError):
            ctx.enable_timing(3)
    finally:
        ctx.close()
