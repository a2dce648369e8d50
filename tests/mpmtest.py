"""Shared test helpers: small seeded scenes and the oracle (CPU restatement).

The oracle library (oracle/liboracle_mpm.so) is the checker; tests call it
on the same inputs as the HIP product path and compare.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_mpm.so")

STICK, SLIP, SEPARATE = 0, 1, 2
COROTATED, NEO_HOOKEAN, HENCKY = 0, 1, 2


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def wall_boxes(dim, N, lo, hi, flags, inset, type_, friction, ghost=3):
    """Axis-aligned wall boxes with the geometry of STANDARD_TESTS_BASE::Add_Walls.

    ``flags`` bit (2*a+s): s=1 is the low side of axis a, s=0 the high side
    (Projects/mpm/STANDARD_TESTS_BASE.cpp:487-511).
    """
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    dx = (hi - lo) / np.asarray(N, float)
    rlo = lo - dx * (ghost * 2 + 1)
    rhi = hi + dx * (ghost * 2 + 1)
    walls = []
    for a in range(dim):
        for s in range(2):
            if flags & (1 << (a * 2 + s)):
                wl, wh = rlo.copy(), rhi.copy()
                if s:
                    wh[a] = lo[a] + inset
                else:
                    wl[a] = hi[a] - inset
                walls.append({"type": type_, "friction": friction, "lo": wl.tolist(), "hi": wh.tolist()})
    return walls


def make_scene(dim, N, prec=8, model=COROTATED, shape=None, ppc_axis=2, seed=1234,
               E=1e4, nu=0.3, rho=1000.0, v0=None, vrand=0.0, brand=0.0, frand=0.0,
               gravity=None, walls=None, dt=1e-4, lo=None, hi=None, jitter=True,
               neo_threshold=0.25, per_particle_E=False):
    """Jittered seeding: each cell split into ppc_axis^d sub-cells, one sample each,
    kept if ``shape(x)`` is true.  Small random perturbations of V, B and F can be
    added so every term of the step is exercised."""
    rng = np.random.default_rng(seed)
    if lo is None:
        lo = [0.0] * dim
    if hi is None:
        hi = [1.0] * dim
    N = [N] * dim if np.isscalar(N) else list(N)
    lo_a, hi_a = np.asarray(lo, float), np.asarray(hi, float)
    dx = (hi_a - lo_a) / np.asarray(N, float)
    sub = np.stack(np.meshgrid(*[np.arange(n * ppc_axis) for n in N], indexing="ij"), -1).reshape(-1, dim)
    u = rng.random(sub.shape) if jitter else np.full(sub.shape, 0.5)
    X = lo_a + (sub + u) * dx / ppc_axis
    if shape is not None:
        X = X[shape(X)]
    n = X.shape[0]
    vol = float(np.prod(dx)) / ppc_axis ** dim
    V = np.zeros((n, dim))
    if v0 is not None:
        V += np.asarray(v0, float)
    if vrand:
        V += vrand * (rng.random((n, dim)) - 0.5)
    F = np.tile(np.eye(dim).reshape(-1), (n, 1))
    if frand:
        F += frand * (rng.random((n, dim * dim)) - 0.5)
    B = np.zeros((n, dim * dim))
    if brand:
        B += brand * (rng.random((n, dim * dim)) - 0.5)
    if per_particle_E:
        Ep = np.exp(rng.uniform(np.log(E / 10), np.log(E * 10), n))
        nup = rng.uniform(0.2, 0.4, n)
        mu, la = lame(Ep, nup)
    else:
        m_, l_ = lame(E, nu)
        mu, la = np.full(n, m_), np.full(n, l_)
    g = [0.0] * dim
    if gravity is None:
        g[1] = -9.8
    else:
        g = list(gravity)
    sc = dict(dim=dim, prec=prec, model=model, N=N, lo=list(lo_a), hi=list(hi_a), gravity=g, dt=dt,
              neo_threshold=neo_threshold, walls=walls or [],
              X=X, V=V, F=F, B=B, mass=np.full(n, rho * vol), volume=np.full(n, vol),
              mu=np.asarray(mu, float), **{"lambda": np.asarray(la, float)}, valid=np.ones(n, np.uint8))
    if prec == 4:
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda"):
            sc[k] = sc[k].astype(np.float32).astype(np.float64)
    return sc


def b_velocity_error(gotB, refB, refV, dx, dim=3):
    """B's error measured by the velocity it carries into the next P2G.  A
    particle adds B D^-1 (x_i - x_p) to the velocity of each stencil node
    (MPM_DRIVER.cpp:263-266), with D^-1 = 4/dx^2 and |x_i - x_p| <= 1.5 dx per
    axis, so an error dB moves a node velocity by at most (6/dx) max_c
    sum_j |dB_cj|; returned relative to the largest particle speed.  This is
    the yardstick when the exact B is ~0 (a body in uniform motion), where
    max|dB| / max|B| only measures rounding noise against noise."""
    d = np.abs(np.asarray(gotB, np.float64) - np.asarray(refB, np.float64)).reshape(-1, dim, dim)
    worst = d.sum(1).max() if d.size else 0.0  # column-major: entry (c, j) at j*dim + c, sum over j
    vmax = max(np.abs(np.asarray(refV, np.float64)).max(), 1e-30)
    return float(6.0 / dx * worst / vmax)


def make_rotating(dim, N, prec=8, model=COROTATED, omega=None, center=None, radius=0.3, E=1e3, nu=0.3,
                  rho=2000.0, seed=1234, dt=1e-3, gravity=None, walls=None):
    """A rotating ball, as the reference's rotating-sphere and rotating-circle
    tests seed it (Projects/mpm/STANDARD_TESTS_3D.cpp:78-91, STANDARD_TESTS_2D.cpp:
    84-99): v = omega x (x - c) with the mean momentum removed, B = (omega x) D_p
    with D_p = dx^2/4 I (STANDARD_TESTS_BASE.cpp:300-303), jittered seeding, no
    gravity unless given."""
    if center is None:
        center = [0.5] * dim
    if omega is None:
        omega = [0.4] if dim == 2 else [0.4, 0.0, 0.0]
    c = np.asarray(center, float)
    ball = lambda X: ((X - c) ** 2).sum(1) < radius ** 2
    sc = make_scene(dim, N, prec=8, model=model, shape=ball, E=E, nu=nu, rho=rho, seed=seed, dt=dt,
                    gravity=gravity if gravity is not None else [0.0] * dim, walls=walls)
    X = sc["X"]
    if dim == 2:
        w = float(omega[0])
        V = np.stack([-w * (X[:, 1] - c[1]), w * (X[:, 0] - c[0])], 1)
        W = np.array([[0.0, -w], [w, 0.0]])
    else:
        w = np.asarray(omega, float)
        V = np.cross(w, X - c)
        W = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    m = sc["mass"]
    V -= (m[:, None] * V).sum(0) / m.sum()
    dx = (np.asarray(sc["hi"]) - np.asarray(sc["lo"])) / np.asarray(sc["N"], float)
    Bm = W * (dx ** 2 / 4.0)[None, :]  # (omega x) D_p, column j scaled by dx_j^2/4
    sc["V"] = V
    sc["B"] = np.tile(Bm.T.reshape(-1), (X.shape[0], 1))  # column-major per particle
    sc["prec"] = prec
    if prec == 4:
        for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda"):
            sc[k] = sc[k].astype(np.float32).astype(np.float64)
    return sc


# ---------------------------------------------------------------- oracle ----
_lib = None


def oracle():
    global _lib
    if _lib is None:
        if not os.path.isfile(ORACLE_SO):
            raise RuntimeError("oracle library missing: run __graft_entry__.build()")
        _lib = ctypes.CDLL(ORACLE_SO)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _walls_arrays(sc):
    walls = sc.get("walls", [])
    nw = len(walls)
    wt = np.array([w["type"] for w in walls] or [0], np.int32)
    wf = np.array([w["friction"] for w in walls] or [0], np.float64)
    wl = np.array([list(w["lo"]) + [0.0] * (3 - sc["dim"]) for w in walls] or [[0, 0, 0]], np.float64)
    wh = np.array([list(w["hi"]) + [0.0] * (3 - sc["dim"]) for w in walls] or [[0, 0, 0]], np.float64)
    return nw, wt, wf, wl, wh


def oracle_run(sc, nsteps, prec=None, want_grid=True):
    """Advance scene ``sc`` by ``nsteps`` with the oracle; returns a dict of final state."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    d = sc["dim"]
    n = sc["X"].shape[0]
    st = {k: np.ascontiguousarray(sc[k], dtype=dt_np).copy() for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda")}
    valid = np.ascontiguousarray(sc["valid"], np.uint8).copy()
    N = np.array(list(sc["N"]) + [1] * (3 - d), np.int32)
    lo = np.array(list(sc["lo"]) + [0] * (3 - d), np.float64)
    hi = np.array(list(sc["hi"]) + [0] * (3 - d), np.float64)
    g = np.array(list(sc["gravity"]) + [0] * (3 - d), np.float64)
    nn = int(np.prod([x + 6 for x in sc["N"]]))
    gm = np.zeros(nn, dt_np) if want_grid else None
    gv = np.zeros((nn, d), dt_np) if want_grid else None
    gs = np.zeros((nn, d), dt_np) if want_grid else None
    gf = np.zeros((nn, d), dt_np) if want_grid else None
    nw, wt, wf, wl, wh = _walls_arrays(sc)
    fn = oracle().oracle_run_f64 if prec == 8 else oracle().oracle_run_f32
    rc = fn(ctypes.c_int(d), _ptr(N), _ptr(lo), _ptr(hi), _ptr(g), ctypes.c_int(sc["model"]),
            ctypes.c_double(sc.get("neo_threshold", 0.25)), ctypes.c_double(sc["dt"]),
            ctypes.c_int(nw), _ptr(wt), _ptr(wf), _ptr(wl), _ptr(wh), ctypes.c_longlong(n),
            _ptr(st["X"]), _ptr(st["V"]), _ptr(st["F"]), _ptr(st["B"]), _ptr(st["mass"]), _ptr(st["volume"]),
            _ptr(st["mu"]), _ptr(st["lambda"]), _ptr(valid), ctypes.c_int(nsteps),
            _ptr(gm), _ptr(gv), _ptr(gs), _ptr(gf))
    assert rc == 0
    out = dict(X=st["X"], V=st["V"], F=st["F"], B=st["B"], valid=valid)
    if want_grid:
        out.update(grid_mass=gm, grid_velocity=gv, grid_velocity_save=gs, grid_velocity_friction_save=gf)
    return out


def oracle_totals(sc, nsteps, prec=None):
    """Conservation totals of ``nsteps`` oracle steps (oracle_run_totals_*): an
    array of (3 nsteps + 1) records of 11 values -- per step the particles at
    its start, the grid after P2G and after the forces; then the particles at
    the end.  Record layout: mass, linear momentum [1:4], angular momentum
    [4:7] (2D: [4]), kinetic energy [7], potential energy [8] = elastic [9] +
    gravity [10]."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    d = sc["dim"]
    n = sc["X"].shape[0]
    st = {k: np.ascontiguousarray(sc[k], dtype=dt_np).copy() for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda")}
    valid = np.ascontiguousarray(sc["valid"], np.uint8).copy()
    N = np.array(list(sc["N"]) + [1] * (3 - d), np.int32)
    lo = np.array(list(sc["lo"]) + [0] * (3 - d), np.float64)
    hi = np.array(list(sc["hi"]) + [0] * (3 - d), np.float64)
    g = np.array(list(sc["gravity"]) + [0] * (3 - d), np.float64)
    out = np.zeros((3 * nsteps + 1, 11), dt_np)
    nw, wt, wf, wl, wh = _walls_arrays(sc)
    fn = oracle().oracle_run_totals_f64 if prec == 8 else oracle().oracle_run_totals_f32
    rc = fn(ctypes.c_int(d), _ptr(N), _ptr(lo), _ptr(hi), _ptr(g), ctypes.c_int(sc["model"]),
            ctypes.c_double(sc.get("neo_threshold", 0.25)), ctypes.c_double(sc["dt"]),
            ctypes.c_int(nw), _ptr(wt), _ptr(wf), _ptr(wl), _ptr(wh), ctypes.c_longlong(n),
            _ptr(st["X"]), _ptr(st["V"]), _ptr(st["F"]), _ptr(st["B"]), _ptr(st["mass"]), _ptr(st["volume"]),
            _ptr(st["mu"]), _ptr(st["lambda"]), _ptr(valid), ctypes.c_int(nsteps), _ptr(out))
    assert rc == 0
    return out.astype(np.float64)


def oracle_run_plastic(sc, nsteps, prec=None):
    """Advance ``nsteps`` with clamped plasticity after every step (sc["plasticity"]: theta_c, theta_s,
    max_hardening, hardening_factor).  Returns the final state with Fp and the hardened mu, lambda."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    d = sc["dim"]
    n = sc["X"].shape[0]
    st = {k: np.ascontiguousarray(sc[k], dtype=dt_np).copy() for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda")}
    valid = np.ascontiguousarray(sc["valid"], np.uint8).copy()
    Fp = np.ascontiguousarray(sc.get("Fp", np.tile(np.eye(d).reshape(-1), (n, 1))), dtype=dt_np).copy()
    N = np.array(list(sc["N"]) + [1] * (3 - d), np.int32)
    lo = np.array(list(sc["lo"]) + [0] * (3 - d), np.float64)
    hi = np.array(list(sc["hi"]) + [0] * (3 - d), np.float64)
    g = np.array(list(sc["gravity"]) + [0] * (3 - d), np.float64)
    pl = sc["plasticity"]
    par = np.array([pl["theta_c"], pl["theta_s"], pl["max_hardening"], pl["hardening_factor"]], np.float64)
    nw, wt, wf, wl, wh = _walls_arrays(sc)
    fn = oracle().oracle_run_plastic_f64 if prec == 8 else oracle().oracle_run_plastic_f32
    rc = fn(ctypes.c_int(d), _ptr(N), _ptr(lo), _ptr(hi), _ptr(g), ctypes.c_int(sc["model"]),
            ctypes.c_double(sc.get("neo_threshold", 0.25)), ctypes.c_double(sc["dt"]),
            ctypes.c_int(nw), _ptr(wt), _ptr(wf), _ptr(wl), _ptr(wh), ctypes.c_longlong(n),
            _ptr(st["X"]), _ptr(st["V"]), _ptr(st["F"]), _ptr(st["B"]), _ptr(st["mass"]), _ptr(st["volume"]),
            _ptr(st["mu"]), _ptr(st["lambda"]), _ptr(valid), ctypes.c_int(nsteps), _ptr(par), _ptr(Fp))
    assert rc == 0
    return dict(X=st["X"], V=st["V"], F=st["F"], B=st["B"], valid=valid, Fp=Fp, mu=st["mu"], **{"lambda": st["lambda"]})


def oracle_run_drucker_prager(sc, nsteps, prec=None):
    """Advance ``nsteps`` with MPM_DRUCKER_PRAGER after every step (sc["plasticity"]: a0, a1, a3, a4,
    cohesion, volume_correction).  Returns the final state with Fp, plastic_def, rho_F, cohesion, vc."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    d = sc["dim"]
    n = sc["X"].shape[0]
    st = {k: np.ascontiguousarray(sc[k], dtype=dt_np).copy() for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda")}
    valid = np.ascontiguousarray(sc["valid"], np.uint8).copy()
    Fp = np.ascontiguousarray(np.tile(np.eye(d).reshape(-1), (n, 1)), dtype=dt_np).copy()
    state = np.zeros(4 * n, dt_np)
    N = np.array(list(sc["N"]) + [1] * (3 - d), np.int32)
    lo = np.array(list(sc["lo"]) + [0] * (3 - d), np.float64)
    hi = np.array(list(sc["hi"]) + [0] * (3 - d), np.float64)
    g = np.array(list(sc["gravity"]) + [0] * (3 - d), np.float64)
    pl = sc["plasticity"]
    par = np.array([pl["a0"], pl["a1"], pl["a3"], pl["a4"], pl["cohesion"],
                    1.0 if pl.get("volume_correction", True) else 0.0], np.float64)
    nw, wt, wf, wl, wh = _walls_arrays(sc)
    fn = oracle().oracle_run_drucker_prager_f64 if prec == 8 else oracle().oracle_run_drucker_prager_f32
    rc = fn(ctypes.c_int(d), _ptr(N), _ptr(lo), _ptr(hi), _ptr(g), ctypes.c_int(sc["model"]),
            ctypes.c_double(sc.get("neo_threshold", 0.25)), ctypes.c_double(sc["dt"]),
            ctypes.c_int(nw), _ptr(wt), _ptr(wf), _ptr(wl), _ptr(wh), ctypes.c_longlong(n),
            _ptr(st["X"]), _ptr(st["V"]), _ptr(st["F"]), _ptr(st["B"]), _ptr(st["mass"]), _ptr(st["volume"]),
            _ptr(st["mu"]), _ptr(st["lambda"]), _ptr(valid), ctypes.c_int(nsteps), _ptr(par), _ptr(Fp), _ptr(state))
    assert rc == 0
    out = dict(X=st["X"], V=st["V"], F=st["F"], B=st["B"], valid=valid, Fp=Fp, mu=st["mu"], **{"lambda": st["lambda"]})
    for i, k in enumerate(("plastic_def", "rho_F", "cohesion", "vc")):
        out[k] = state[i * n:(i + 1) * n]
    return out


def _limits_array(lim):
    """(strong_cfl, sound_cfl, cfl_F, cfl_sound) from a dict, or None."""
    if not lim:
        return None
    return np.array([1.0 if lim.get("strong_cfl") else 0.0, 1.0 if lim.get("sound_cfl") else 0.0,
                     lim.get("cfl_F", 0.5), lim.get("cfl_sound", 0.9)], np.float64)


def oracle_run_limited(sc, nsteps, lim, prec=None):
    """``nsteps`` fixed-dt steps with Reduce_Dt (lim: strong_cfl, sound_cfl, cfl, cfl_F, cfl_sound, min_dt).

    Returns (final state dict, dts taken)."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    d = sc["dim"]
    n = sc["X"].shape[0]
    st = {k: np.ascontiguousarray(sc[k], dtype=dt_np).copy() for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda")}
    valid = np.ascontiguousarray(sc["valid"], np.uint8).copy()
    N = np.array(list(sc["N"]) + [1] * (3 - d), np.int32)
    lo = np.array(list(sc["lo"]) + [0] * (3 - d), np.float64)
    hi = np.array(list(sc["hi"]) + [0] * (3 - d), np.float64)
    g = np.array(list(sc["gravity"]) + [0] * (3 - d), np.float64)
    nw, wt, wf, wl, wh = _walls_arrays(sc)
    la_ = _limits_array(lim)
    dts = np.zeros(max(nsteps, 1), np.float64)
    fn = oracle().oracle_run_limited_f64 if prec == 8 else oracle().oracle_run_limited_f32
    rc = fn(ctypes.c_int(d), _ptr(N), _ptr(lo), _ptr(hi), _ptr(g), ctypes.c_int(sc["model"]),
            ctypes.c_double(sc.get("neo_threshold", 0.25)), ctypes.c_double(sc["dt"]),
            ctypes.c_int(nw), _ptr(wt), _ptr(wf), _ptr(wl), _ptr(wh), ctypes.c_longlong(n),
            _ptr(st["X"]), _ptr(st["V"]), _ptr(st["F"]), _ptr(st["B"]), _ptr(st["mass"]), _ptr(st["volume"]),
            _ptr(st["mu"]), _ptr(st["lambda"]), _ptr(valid), ctypes.c_int(nsteps), _ptr(la_),
            ctypes.c_double(lim.get("cfl", 1.0)), ctypes.c_double(lim.get("min_dt", 0.0)), _ptr(dts))
    assert rc == 0
    return dict(X=st["X"], V=st["V"], F=st["F"], B=st["B"], valid=valid), dts[:nsteps].copy()


FREE, SLIP_BC, NOSLIP_BC, SEPARATE_BC = 0, 1, 2, 3


def oracle_run_reflect(sc, nsteps, prec=None):
    """``nsteps`` steps with reflection boundary conditions (sc["reflection"]: full, friction, sides)."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    d = sc["dim"]
    n = sc["X"].shape[0]
    st = {k: np.ascontiguousarray(sc[k], dtype=dt_np).copy() for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda")}
    valid = np.ascontiguousarray(sc["valid"], np.uint8).copy()
    N = np.array(list(sc["N"]) + [1] * (3 - d), np.int32)
    lo = np.array(list(sc["lo"]) + [0] * (3 - d), np.float64)
    hi = np.array(list(sc["hi"]) + [0] * (3 - d), np.float64)
    g = np.array(list(sc["gravity"]) + [0] * (3 - d), np.float64)
    nw, wt, wf, wl, wh = _walls_arrays(sc)
    r = sc["reflection"]
    refl = np.array([1.0 if r.get("full") else 0.0, r.get("friction", 0.0)] + [float(t) for t in r["sides"]], np.float64)
    fn = oracle().oracle_run_reflect_f64 if prec == 8 else oracle().oracle_run_reflect_f32
    rc = fn(ctypes.c_int(d), _ptr(N), _ptr(lo), _ptr(hi), _ptr(g), ctypes.c_int(sc["model"]),
            ctypes.c_double(sc.get("neo_threshold", 0.25)), ctypes.c_double(sc["dt"]),
            ctypes.c_int(nw), _ptr(wt), _ptr(wf), _ptr(wl), _ptr(wh), ctypes.c_longlong(n),
            _ptr(st["X"]), _ptr(st["V"]), _ptr(st["F"]), _ptr(st["B"]), _ptr(st["mass"]), _ptr(st["volume"]),
            _ptr(st["mu"]), _ptr(st["lambda"]), _ptr(valid), ctypes.c_int(nsteps), _ptr(refl))
    assert rc == 0
    return dict(X=st["X"], V=st["V"], F=st["F"], B=st["B"], valid=valid)


def oracle_sound_speeds(sc, prec=None, F=None):
    """Sound_Speed of every particle of ``sc`` (or of the given F) from the singular values of F."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    n = sc["X"].shape[0]
    arr = {k: np.ascontiguousarray(sc[k] if k != "F" or F is None else F, dtype=dt_np)
           for k in ("F", "mass", "volume", "mu", "lambda")}
    out = np.zeros(n, dt_np)
    fn = oracle().oracle_sound_speeds_f64 if prec == 8 else oracle().oracle_sound_speeds_f32
    rc = fn(ctypes.c_int(sc["dim"]), ctypes.c_int(sc["model"]), ctypes.c_double(sc.get("neo_threshold", 0.25)),
            ctypes.c_longlong(n), _ptr(arr["F"]), _ptr(arr["mass"]), _ptr(arr["volume"]), _ptr(arr["mu"]),
            _ptr(arr["lambda"]), _ptr(out))
    assert rc == 0
    return out


def oracle_frames(sc, nframes, frame_dt, cfl, min_dt, max_dt, prec=None, max_sub=100000, lim=None):
    """Advance ``nframes`` frames with the adaptive substep loop of the oracle
    (with Reduce_Dt when ``lim`` turns the strong or sound-speed CFL on).

    Returns (final state dict, array of substep dts)."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    d = sc["dim"]
    n = sc["X"].shape[0]
    st = {k: np.ascontiguousarray(sc[k], dtype=dt_np).copy() for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lambda")}
    valid = np.ascontiguousarray(sc["valid"], np.uint8).copy()
    N = np.array(list(sc["N"]) + [1] * (3 - d), np.int32)
    lo = np.array(list(sc["lo"]) + [0] * (3 - d), np.float64)
    hi = np.array(list(sc["hi"]) + [0] * (3 - d), np.float64)
    g = np.array(list(sc["gravity"]) + [0] * (3 - d), np.float64)
    nw, wt, wf, wl, wh = _walls_arrays(sc)
    dts = np.zeros(max_sub, np.float64)
    fn = oracle().oracle_advance_frames_f64 if prec == 8 else oracle().oracle_advance_frames_f32
    cnt = fn(ctypes.c_int(d), _ptr(N), _ptr(lo), _ptr(hi), _ptr(g), ctypes.c_int(sc["model"]),
             ctypes.c_double(sc.get("neo_threshold", 0.25)), ctypes.c_int(nw), _ptr(wt), _ptr(wf), _ptr(wl), _ptr(wh),
             ctypes.c_longlong(n), _ptr(st["X"]), _ptr(st["V"]), _ptr(st["F"]), _ptr(st["B"]), _ptr(st["mass"]),
             _ptr(st["volume"]), _ptr(st["mu"]), _ptr(st["lambda"]), _ptr(valid), ctypes.c_int(nframes),
             ctypes.c_double(frame_dt), ctypes.c_double(cfl), ctypes.c_double(min_dt), ctypes.c_double(max_dt),
             _ptr(dts), ctypes.c_int(max_sub), _ptr(_limits_array(lim)))
    assert cnt >= 0, "substep limit exceeded"
    return dict(X=st["X"], V=st["V"], F=st["F"], B=st["B"], valid=valid), dts[:cnt].copy()


def oracle_grid_v_bound(sc, prec=None):
    """Grid_V_Upper_Bound of the particles of ``sc`` (max |v| + 6 sqrt(d)/min dx |B|_F over valid particles)."""
    prec = prec or sc["prec"]
    dt_np = np.float64 if prec == 8 else np.float32
    d = sc["dim"]
    V = np.ascontiguousarray(sc["V"], dtype=dt_np)
    B = np.ascontiguousarray(sc["B"], dtype=dt_np)
    valid = np.ascontiguousarray(sc["valid"], np.uint8)
    N = np.array(list(sc["N"]) + [1] * (3 - d), np.int32)
    lo = np.array(list(sc["lo"]) + [0] * (3 - d), np.float64)
    hi = np.array(list(sc["hi"]) + [0] * (3 - d), np.float64)
    out = np.zeros(1, np.float64)
    fn = oracle().oracle_grid_v_bound_f64 if prec == 8 else oracle().oracle_grid_v_bound_f32
    rc = fn(ctypes.c_int(d), _ptr(N), _ptr(lo), _ptr(hi), ctypes.c_longlong(V.shape[0]), _ptr(V), _ptr(B),
            _ptr(valid), _ptr(out))
    assert rc == 0
    return float(out[0])


def oracle_svd(A, dim, prec=8):
    dt_np = np.float64 if prec == 8 else np.float32
    A = np.ascontiguousarray(A, dt_np).reshape(-1, dim * dim)
    n = A.shape[0]
    U = np.zeros_like(A)
    V = np.zeros_like(A)
    S = np.zeros((n, dim), dt_np)
    it = np.zeros(n, np.int32)
    fn = oracle().oracle_svd_f64 if prec == 8 else oracle().oracle_svd_f32
    assert fn(ctypes.c_int(dim), ctypes.c_longlong(n), _ptr(A), _ptr(U), _ptr(S), _ptr(V), _ptr(it)) == 0
    return U, S, V, it


def oracle_pstrain(S, mu, la, dim, model, thr=0.25, prec=8):
    dt_np = np.float64 if prec == 8 else np.float32
    S = np.ascontiguousarray(S, dt_np).reshape(-1, dim)
    n = S.shape[0]
    mu = np.ascontiguousarray(np.broadcast_to(mu, (n,)), dt_np)
    la = np.ascontiguousarray(np.broadcast_to(la, (n,)), dt_np)
    P = np.zeros_like(S)
    fn = oracle().oracle_pstrain_f64 if prec == 8 else oracle().oracle_pstrain_f32
    assert fn(ctypes.c_int(dim), ctypes.c_int(model), ctypes.c_double(thr), ctypes.c_longlong(n),
              _ptr(S), _ptr(mu), _ptr(la), _ptr(P)) == 0
    return P


def oracle_stress(F, vol, mu, la, dim, model, thr=0.25, prec=8):
    dt_np = np.float64 if prec == 8 else np.float32
    F = np.ascontiguousarray(F, dt_np).reshape(-1, dim * dim)
    n = F.shape[0]
    b = lambda a: np.ascontiguousarray(np.broadcast_to(a, (n,)), dt_np)
    vol, mu, la = b(vol), b(mu), b(la)
    A = np.zeros_like(F)
    fn = oracle().oracle_stress_f64 if prec == 8 else oracle().oracle_stress_f32
    assert fn(ctypes.c_int(dim), ctypes.c_int(model), ctypes.c_double(thr), ctypes.c_longlong(n),
              _ptr(F), _ptr(vol), _ptr(mu), _ptr(la), _ptr(A)) == 0
    return A


def oracle_weights(X, dim, N, lo, hi, prec=8):
    dt_np = np.float64 if prec == 8 else np.float32
    X = np.ascontiguousarray(X, dt_np).reshape(-1, dim)
    n = X.shape[0]
    Na = np.array([N] * 3, np.int32)
    loa = np.array([lo] * 3, np.float64)
    hia = np.array([hi] * 3, np.float64)
    base = np.zeros((n, dim), np.int32)
    w = np.zeros((n, 3, dim), dt_np)
    dw = np.zeros((n, 3, dim), dt_np)
    fn = oracle().oracle_weights_f64 if prec == 8 else oracle().oracle_weights_f32
    assert fn(ctypes.c_int(dim), _ptr(Na), _ptr(loa), _ptr(hia), ctypes.c_longlong(n), _ptr(X),
              _ptr(base), _ptr(w), _ptr(dw)) == 0
    return base, w, dw
