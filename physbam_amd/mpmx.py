"""ctypes binding of libmpmx.so (include/mpmx.h and include/mpmx_scene.h).

The binding is thin: every call goes to the C ABI and raises MpmxError on a
non-zero status.  There is no CPU fallback; importing works without a GPU,
but creating a context needs the HIP device.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpmx.so")

FP32, FP64 = 4, 8
FIXED_COROTATED, NEO_HOOKEAN, HENCKY = 0, 1, 2
STICK, SLIP, SEPARATE = 0, 1, 2
SHAPE_BOX, SHAPE_SPHERE = 0, 1
# mpmx_totals: mass, linear[3], angular[3], kinetic, potential, elastic, gravity, count
TOTALS_LEN = 12

# Every exported symbol of the C ABI, checked by the CPU test suite.
EXPORTS = (
    "mpmx_last_error", "mpmx_create", "mpmx_destroy", "mpmx_create_batch", "mpmx_scene_count",
    "mpmx_set_batch_particles", "mpmx_set_particles", "mpmx_get_particles", "mpmx_get_material",
    "mpmx_particle_count", "mpmx_set_constitutive", "mpmx_set_gravity", "mpmx_add_collision_box",
    "mpmx_clear_collision_objects", "mpmx_step", "mpmx_step_async", "mpmx_synchronize", "mpmx_time",
    "mpmx_step_stage", "mpmx_stage_name",
    "mpmx_enable_grid_stats", "mpmx_grid_stats", "mpmx_particle_stats", "mpmx_particle_totals",
    "mpmx_grid_totals", "mpmx_grid_node_count",
    "mpmx_get_grid", "mpmx_set_reorder_interval", "mpmx_set_rebin_policy", "mpmx_bin_stats",
    "mpmx_grid_v_upper_bound", "mpmx_compute_dt", "mpmx_advance_frames",
    "mpmx_set_dt_limits", "mpmx_max_sound_speed", "mpmx_last_dt", "mpmx_set_reflection",
    "mpmx_write_particles_file", "mpmx_read_particles_file", "mpmx_write_output_files", "mpmx_read_output_files",
    "mpmx_write_restart_file", "mpmx_read_restart_file",
    "mpmx_set_plasticity", "mpmx_get_plastic_state", "mpmx_set_drucker_prager", "mpmx_get_drucker_prager_state",
    "mpmx_tape_begin", "mpmx_backward",
    "mpmx_enable_timing", "mpmx_reset_timing",
    "mpmx_phase_time", "mpmx_stress", "mpmx_svd", "mpmx_sort_keys", "mpmx_sort_keys_noalloc", "mpmx_seed_particles",
    "mpmx_set_launch_config", "mpmx_get_launch_config", "mpmx_autotune",
    "mpmx_random_boxes", "mpmx_seed_particles_poisson",
)


STAGES = 8


def stage_name(stage):
    """Name of a stage of mpmx_step_stage ("simulated-particles" ... "plasticity-hardening"), or None."""
    r = lib().mpmx_stage_name(int(stage))
    return None if r is None else r.decode()


class MpmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mpmx error {code}: {msg}")
        self.code = code


class GridDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("precision", ctypes.c_int), ("cells", ctypes.c_int * 3),
                ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3)]


ARRAY_SCALAR, ARRAY_VECTOR, ARRAY_VECTOR3, ARRAY_MATRIX, ARRAY_BOOL = 0, 1, 2, 3, 4
BC_FREE, BC_SLIP, BC_NOSLIP, BC_SEPARATE = 0, 1, 2, 3


class ParticleArray(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("kind", ctypes.c_int), ("data", ctypes.c_void_p)]


class SeedDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("cells", ctypes.c_int * 3), ("lo", ctypes.c_double * 3),
                ("hi", ctypes.c_double * 3), ("ppc_axis", ctypes.c_int), ("seed", ctypes.c_uint32),
                ("jitter", ctypes.c_int), ("nshapes", ctypes.c_int),
                ("shape_type", ctypes.c_void_p), ("shape_a", ctypes.c_void_p), ("shape_b", ctypes.c_void_p),
                ("density", ctypes.c_void_p), ("youngs", ctypes.c_void_p), ("poisson", ctypes.c_void_p),
                ("velocity", ctypes.c_void_p)]


_lib = None


class LaunchConfig(ctypes.Structure):
    """mpmx_launch_config: graph (0/1), tile_workgroups (0 = one per tile up to 8192), tile_threads (0/128/256)."""
    _fields_ = [("graph", ctypes.c_int), ("tile_workgroups", ctypes.c_int), ("tile_threads", ctypes.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise MpmxError(-5, f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, c_int, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        L.mpmx_last_error.restype = ctypes.c_char_p
        L.mpmx_create.argtypes = [ctypes.POINTER(GridDesc), ctypes.POINTER(vp)]
        L.mpmx_destroy.argtypes = [vp]
        L.mpmx_create_batch.argtypes = [ctypes.POINTER(GridDesc), c_int, ctypes.POINTER(vp)]
        L.mpmx_scene_count.argtypes = [vp]
        L.mpmx_set_batch_particles.argtypes = [vp, c_int, vp] + [vp] * 9
        L.mpmx_set_particles.argtypes = [vp, i64] + [vp] * 9
        L.mpmx_get_particles.argtypes = [vp] + [vp] * 5
        L.mpmx_particle_count.argtypes = [vp]
        L.mpmx_particle_count.restype = i64
        L.mpmx_set_constitutive.argtypes = [vp, c_int, dbl]
        L.mpmx_set_gravity.argtypes = [vp, vp]
        L.mpmx_add_collision_box.argtypes = [vp, c_int, dbl, vp, vp]
        L.mpmx_clear_collision_objects.argtypes = [vp]
        L.mpmx_step.argtypes = [vp, dbl, c_int]
        L.mpmx_step_async.argtypes = [vp, dbl, c_int]
        L.mpmx_step_stage.argtypes = [vp, dbl, c_int]
        L.mpmx_stage_name.argtypes = [c_int]
        L.mpmx_stage_name.restype = ctypes.c_char_p
        L.mpmx_synchronize.argtypes = [vp]
        L.mpmx_time.argtypes = [vp]
        L.mpmx_time.restype = dbl
        L.mpmx_enable_grid_stats.argtypes = [vp, c_int]
        L.mpmx_grid_stats.argtypes = [vp, vp, vp]
        L.mpmx_particle_stats.argtypes = [vp, vp, vp, vp]
        L.mpmx_particle_totals.argtypes = [vp, vp]
        L.mpmx_grid_totals.argtypes = [vp, c_int, vp]
        L.mpmx_grid_node_count.argtypes = [vp]
        L.mpmx_grid_node_count.restype = i64
        L.mpmx_get_grid.argtypes = [vp, vp, vp]
        L.mpmx_set_reorder_interval.argtypes = [vp, c_int]
        L.mpmx_set_rebin_policy.argtypes = [vp, dbl, c_int]
        L.mpmx_bin_stats.argtypes = [vp, vp, vp, vp]
        L.mpmx_grid_v_upper_bound.argtypes = [vp, vp]
        L.mpmx_compute_dt.argtypes = [vp, dbl, dbl, vp]
        L.mpmx_advance_frames.argtypes = [vp, c_int, dbl, dbl, dbl, dbl, vp, i64, vp]
        L.mpmx_set_plasticity.argtypes = [vp, c_int, dbl, dbl, dbl, dbl]
        L.mpmx_get_plastic_state.argtypes = [vp, vp, vp, vp]
        L.mpmx_get_material.argtypes = [vp, vp, vp, vp, vp]
        L.mpmx_write_particles_file.argtypes = [ctypes.c_char_p, c_int, c_int, i64, ctypes.POINTER(ParticleArray), c_int]
        L.mpmx_read_particles_file.argtypes = [ctypes.c_char_p, c_int, c_int, ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(ParticleArray), c_int, vp]
        L.mpmx_write_output_files.argtypes = [vp, ctypes.c_char_p]
        L.mpmx_write_restart_file.argtypes = [ctypes.c_char_p, c_int, dbl]
        L.mpmx_read_restart_file.argtypes = [ctypes.c_char_p, vp]
        L.mpmx_read_output_files.argtypes = [vp, ctypes.c_char_p]
        L.mpmx_set_reflection.argtypes = [vp, c_int, c_int, dbl, vp]
        L.mpmx_set_dt_limits.argtypes = [vp, c_int, c_int, dbl, dbl, dbl, dbl]
        L.mpmx_max_sound_speed.argtypes = [vp, vp]
        L.mpmx_last_dt.argtypes = [vp, vp]
        L.mpmx_set_drucker_prager.argtypes = [vp, c_int, dbl, dbl, dbl, dbl, dbl, c_int]
        L.mpmx_get_drucker_prager_state.argtypes = [vp, vp, vp, vp, vp]
        L.mpmx_tape_begin.argtypes = [vp, c_int]
        L.mpmx_backward.argtypes = [vp] + [vp] * 10
        L.mpmx_enable_timing.argtypes = [vp, c_int]
        L.mpmx_reset_timing.argtypes = [vp]
        L.mpmx_phase_time.argtypes = [vp, ctypes.c_char_p, vp, vp]
        L.mpmx_stress.argtypes = [vp, i64] + [vp] * 5
        L.mpmx_svd.argtypes = [vp, i64] + [vp] * 4
        L.mpmx_sort_keys.argtypes = [vp, i64, vp, c_int, vp, vp]
        L.mpmx_sort_keys_noalloc.argtypes = [vp, i64, vp, c_int, vp, vp]
        L.mpmx_set_launch_config.argtypes = [vp, ctypes.POINTER(LaunchConfig)]
        L.mpmx_get_launch_config.argtypes = [vp, ctypes.POINTER(LaunchConfig)]
        L.mpmx_autotune.argtypes = [vp, dbl, c_int, ctypes.POINTER(LaunchConfig), vp, c_int, ctypes.POINTER(c_int)]
        L.mpmx_seed_particles.argtypes = [ctypes.POINTER(SeedDesc), c_int, i64] + [vp] * 9
        L.mpmx_seed_particles.restype = i64
        L.mpmx_seed_particles_poisson.argtypes = [ctypes.POINTER(SeedDesc), c_int, dbl, i64] + [vp] * 9
        L.mpmx_seed_particles_poisson.restype = i64
        L.mpmx_random_boxes.argtypes = [ctypes.c_uint32, c_int, dbl, dbl, dbl, dbl, dbl, dbl, dbl, dbl, c_int,
                                        c_int, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise MpmxError(rc, lib().mpmx_last_error().decode())


def _array_shape(kind, dim):
    return {ARRAY_SCALAR: (), ARRAY_VECTOR: (dim,), ARRAY_VECTOR3: (3,), ARRAY_MATRIX: (dim * dim,),
            ARRAY_BOOL: ()}[kind]


def write_particles_file(path, dim, precision, arrays):
    """PARTICLES<TV>::Write of ``arrays`` = {name: (kind, array)} to ``path`` (gzip if it ends in .gz)."""
    dt = np.float64 if precision == FP64 else np.float32
    keep, descs, n = [], [], None
    for name, (kind, a) in arrays.items():
        a = np.ascontiguousarray(a, np.uint8 if kind == ARRAY_BOOL else dt)
        m = a.shape[0]
        n = m if n is None else n
        width = int(np.prod(_array_shape(kind, dim)))
        if m != n or a.size != m * width:
            raise MpmxError(-1, f"array {name} has the wrong shape")
        keep.append(a)
        descs.append(ParticleArray(name.encode(), kind, a.ctypes.data))
    arr = (ParticleArray * len(descs))(*descs)
    _check(lib().mpmx_write_particles_file(os.fsencode(path), dim, precision, n or 0, arr, len(descs)))


def read_particles_file(path, dim, precision, kinds):
    """PARTICLES<TV>::Read of the arrays named in ``kinds`` = {name: kind}; missing ones are left out."""
    dt = np.float64 if precision == FP64 else np.float32
    n = ctypes.c_int64(0)
    _check(lib().mpmx_read_particles_file(os.fsencode(path), dim, precision, ctypes.byref(n), None, 0, None))
    out = {k: np.zeros((n.value,) + _array_shape(kind, dim), np.uint8 if kind == ARRAY_BOOL else dt)
           for k, kind in kinds.items()}
    descs = [ParticleArray(k.encode(), kinds[k], out[k].ctypes.data) for k in kinds]
    arr = (ParticleArray * len(descs))(*descs)
    found = np.zeros(len(descs), np.int32)
    _check(lib().mpmx_read_particles_file(os.fsencode(path), dim, precision, ctypes.byref(n), arr, len(descs),
                                          found.ctypes.data))
    return {k: out[k] for i, k in enumerate(kinds) if found[i]}


def write_restart_file(path, precision, time):
    _check(lib().mpmx_write_restart_file(os.fsencode(path), precision, float(time)))


def read_restart_file(path):
    out = np.zeros(1)
    _check(lib().mpmx_read_restart_file(os.fsencode(path), out.ctypes.data))
    return float(out[0])


class Context:
    """One MPM scene on the current GPU (MPM_EXAMPLE + MPM_DRIVER state)."""

    def __init__(self, dim, cells, lo, hi, precision=FP32):
        self.dim = dim
        self.precision = precision
        self.dtype = np.float32 if precision == FP32 else np.float64
        cells = [cells] * dim if np.isscalar(cells) else list(cells)
        self.cells = cells
        d = GridDesc()
        d.dim, d.precision = dim, precision
        for a in range(3):
            d.cells[a] = cells[a] if a < dim else 1
            d.lo[a] = lo[a] if a < dim else 0.0
            d.hi[a] = hi[a] if a < dim else 1.0
        h = ctypes.c_void_p()
        _check(lib().mpmx_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    @classmethod
    def batch(cls, dim, grids, precision=FP32):
        """A batch of independent scenes; grids: list of (cells, lo, hi)."""
        self = cls.__new__(cls)
        self.dim, self.precision = dim, precision
        self.dtype = np.float32 if precision == FP32 else np.float64
        descs = (GridDesc * len(grids))()
        for s, (cells, lo, hi) in enumerate(grids):
            cells = [cells] * dim if np.isscalar(cells) else list(cells)
            descs[s].dim, descs[s].precision = dim, precision
            for a in range(3):
                descs[s].cells[a] = cells[a] if a < dim else 1
                descs[s].lo[a] = lo[a] if a < dim else 0.0
                descs[s].hi[a] = hi[a] if a < dim else 1.0
        h = ctypes.c_void_p()
        _check(lib().mpmx_create_batch(descs, len(grids), ctypes.byref(h)))
        self._h = h
        self.cells = None
        return self

    def set_batch_particles(self, counts, X, V, F, B, mass, volume, mu, lam, valid=None):
        """Particles of all scenes concatenated in scene order; counts[s] for scene s."""
        d = self.dim
        n = int(np.asarray(X).shape[0])
        counts = np.ascontiguousarray(counts, np.int64)
        assert counts.sum() == n
        arrs = [self._arr(X, (n, d)), self._arr(V, (n, d)), self._arr(F, (n, d * d)), self._arr(B, (n, d * d)),
                self._arr(mass, (n,)), self._arr(volume, (n,)), self._arr(mu, (n,)), self._arr(lam, (n,))]
        v = np.ones(n, np.uint8) if valid is None else np.ascontiguousarray(valid, np.uint8)
        _check(lib().mpmx_set_batch_particles(self._h, len(counts), _p(counts), *[_p(a) for a in arrs], _p(v)))
        self.n = n

    def close(self):
        if getattr(self, "_h", None):
            lib().mpmx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _arr(self, a, shape):
        return np.ascontiguousarray(np.asarray(a, dtype=self.dtype).reshape(shape))

    def set_particles(self, X, V, F, B, mass, volume, mu, lam, valid=None):
        d = self.dim
        n = int(np.asarray(X).shape[0])
        arrs = [self._arr(X, (n, d)), self._arr(V, (n, d)), self._arr(F, (n, d * d)), self._arr(B, (n, d * d)),
                self._arr(mass, (n,)), self._arr(volume, (n,)), self._arr(mu, (n,)), self._arr(lam, (n,))]
        v = np.ones(n, np.uint8) if valid is None else np.ascontiguousarray(valid, np.uint8)
        _check(lib().mpmx_set_particles(self._h, n, *[_p(a) for a in arrs], _p(v)))
        self.n = n

    def get_particles(self):
        n, d = self.n, self.dim
        X = np.zeros((n, d), self.dtype)
        V = np.zeros((n, d), self.dtype)
        F = np.zeros((n, d * d), self.dtype)
        B = np.zeros((n, d * d), self.dtype)
        valid = np.zeros(n, np.uint8)
        _check(lib().mpmx_get_particles(self._h, _p(X), _p(V), _p(F), _p(B), _p(valid)))
        return dict(X=X, V=V, F=F, B=B, valid=valid)

    def set_constitutive(self, model, neo_threshold=0.25):
        _check(lib().mpmx_set_constitutive(self._h, model, neo_threshold))

    def set_gravity(self, g):
        g = np.ascontiguousarray(list(g) + [0.0] * (3 - len(g)), np.float64)
        _check(lib().mpmx_set_gravity(self._h, _p(g)))

    def add_collision_box(self, type_, friction, lo, hi):
        lo = np.ascontiguousarray(list(lo) + [0.0] * (3 - len(lo)), np.float64)
        hi = np.ascontiguousarray(list(hi) + [0.0] * (3 - len(hi)), np.float64)
        _check(lib().mpmx_add_collision_box(self._h, type_, friction, _p(lo), _p(hi)))

    def clear_collision_objects(self):
        _check(lib().mpmx_clear_collision_objects(self._h))

    def step(self, dt, nsteps=1):
        _check(lib().mpmx_step(self._h, dt, nsteps))

    def step_async(self, dt, nsteps=1):
        _check(lib().mpmx_step_async(self._h, dt, nsteps))

    def step_stage(self, dt, stage):
        """One named step of MPM_DRIVER::Advance_One_Time_Step (stage 0..7, see stage_name)."""
        _check(lib().mpmx_step_stage(self._h, dt, int(stage)))

    def synchronize(self):
        _check(lib().mpmx_synchronize(self._h))

    @property
    def time(self):
        return lib().mpmx_time(self._h)

    def enable_grid_stats(self, on=True):
        _check(lib().mpmx_enable_grid_stats(self._h, 1 if on else 0))

    def grid_stats(self):
        m = np.zeros(1)
        p = np.zeros(3)
        _check(lib().mpmx_grid_stats(self._h, _p(m), _p(p)))
        return float(m[0]), p[: self.dim].copy()

    def particle_stats(self):
        m = np.zeros(1)
        p = np.zeros(3)
        c = np.zeros(1, np.int64)
        _check(lib().mpmx_particle_stats(self._h, _p(m), _p(p), _p(c)))
        return float(m[0]), p[: self.dim].copy(), int(c[0])

    def particle_totals(self):
        """FP64 particle totals (mpmx_totals as 12 doubles: mass, linear[3],
        angular[3], kinetic, potential, elastic, gravity, count)."""
        out = np.zeros(TOTALS_LEN)
        _check(lib().mpmx_particle_totals(self._h, _p(out)))
        return out

    def grid_totals(self, which):
        """FP64 grid totals of the last step: which 0 after P2G, 1 after the grid
        update (needs enable_grid_stats)."""
        out = np.zeros(TOTALS_LEN)
        _check(lib().mpmx_grid_totals(self._h, int(which), _p(out)))
        return out

    def get_grid(self):
        nn = lib().mpmx_grid_node_count(self._h)
        m = np.zeros(nn, self.dtype)
        v = np.zeros((nn, self.dim), self.dtype)
        _check(lib().mpmx_get_grid(self._h, _p(m), _p(v)))
        return m, v

    def set_launch_config(self, graph=False, tile_workgroups=0, tile_threads=0):
        c = LaunchConfig(1 if graph else 0, int(tile_workgroups), int(tile_threads))
        _check(lib().mpmx_set_launch_config(self._h, ctypes.byref(c)))

    def launch_config(self):
        c = LaunchConfig()
        _check(lib().mpmx_get_launch_config(self._h, ctypes.byref(c)))
        return {"graph": bool(c.graph), "tile_workgroups": int(c.tile_workgroups), "tile_threads": int(c.tile_threads)}

    def autotune(self, dt, trial_steps=16):
        """mpmx_autotune: returns (chosen config dict, list of (config dict, ms per step) per candidate)."""
        c = LaunchConfig()
        times = np.zeros(32)
        nt = ctypes.c_int(0)
        _check(lib().mpmx_autotune(self._h, float(dt), int(trial_steps), ctypes.byref(c), _p(times), 32,
                                   ctypes.byref(nt)))
        ch = {"graph": bool(c.graph), "tile_workgroups": int(c.tile_workgroups), "tile_threads": int(c.tile_threads)}
        return ch, [float(t) for t in times[:nt.value]]

    def set_reorder_interval(self, k):
        _check(lib().mpmx_set_reorder_interval(self._h, k))

    def set_rebin_policy(self, mover_fraction=1.0 / 128, max_interval=0):
        _check(lib().mpmx_set_rebin_policy(self._h, float(mover_fraction), int(max_interval)))

    def bin_stats(self, with_slow=False):
        """(bin rebuilds so far, movers in the last step[, FP64-SVD stress particles in the last step])."""
        r = np.zeros(3, np.int64)
        _check(lib().mpmx_bin_stats(self._h, _p(r[0:1]), _p(r[1:2]), _p(r[2:3])))
        return (int(r[0]), int(r[1]), int(r[2])) if with_slow else (int(r[0]), int(r[1]))

    def grid_v_upper_bound(self):
        """Per-scene max over valid particles of |v| + 6 sqrt(dim) min(1/dx) |B|_F (float64 array)."""
        out = np.zeros(lib().mpmx_scene_count(self._h))
        _check(lib().mpmx_grid_v_upper_bound(self._h, _p(out)))
        return out

    def compute_dt(self, cfl=1.0, max_dt=1.0 / 24):
        """Per-scene CFL time step (float64 array)."""
        out = np.zeros(lib().mpmx_scene_count(self._h))
        _check(lib().mpmx_compute_dt(self._h, float(cfl), float(max_dt), _p(out)))
        return out

    def advance_frames(self, nframes, frame_dt=1.0 / 24, cfl=1.0, min_dt=0.0, max_dt=None, capacity=1 << 20):
        """Advance ``nframes`` frames with adaptive substeps; returns the substep dts."""
        max_dt = frame_dt if max_dt is None else max_dt
        dts = np.zeros(capacity)
        cnt = np.zeros(1, np.int64)
        _check(lib().mpmx_advance_frames(self._h, int(nframes), float(frame_dt), float(cfl), float(min_dt),
                                         float(max_dt), _p(dts), ctypes.c_int64(capacity), _p(cnt)))
        return dts[:min(int(cnt[0]), capacity)].copy()

    def get_material(self):
        """dict mass, volume, mu, lambda of the particles (user order)."""
        out = {k: np.zeros(self.n, self.dtype) for k in ("mass", "volume", "mu", "lambda")}
        _check(lib().mpmx_get_material(self._h, *(_p(out[k]) for k in ("mass", "volume", "mu", "lambda"))))
        return out

    def write_output_files(self, frame_dir):
        """MPM_EXAMPLE::Write_Output_Files' particle and restart files into ``frame_dir``."""
        _check(lib().mpmx_write_output_files(self._h, os.fsencode(frame_dir)))

    def read_output_files(self, frame_dir):
        """MPM_EXAMPLE::Read_Output_Files: particles and time from ``frame_dir``."""
        _check(lib().mpmx_read_output_files(self._h, os.fsencode(frame_dir)))

    def set_reflection(self, sides, full=False, friction=0.0, enable=True):
        """Reflection boundary conditions; sides: 2*dim types (BC_FREE, BC_SLIP, BC_NOSLIP, BC_SEPARATE)."""
        t = np.ascontiguousarray(sides, np.int32)
        _check(lib().mpmx_set_reflection(self._h, 1 if enable else 0, 1 if full else 0, float(friction), _p(t)))

    def set_dt_limits(self, strong_cfl=False, sound_cfl=False, cfl=1.0, cfl_F=0.5, cfl_sound=0.9, min_dt=0.0):
        """Strong and sound-speed CFL (MPM_DRIVER::Reduce_Dt) inside every step."""
        _check(lib().mpmx_set_dt_limits(self._h, 1 if strong_cfl else 0, 1 if sound_cfl else 0, float(cfl),
                                        float(cfl_F), float(cfl_sound), float(min_dt)))

    def max_sound_speed(self):
        out = np.zeros(1)
        _check(lib().mpmx_max_sound_speed(self._h, _p(out)))
        return float(out[0])

    @property
    def last_dt(self):
        out = np.zeros(1)
        _check(lib().mpmx_last_dt(self._h, _p(out)))
        return float(out[0])

    def set_plasticity(self, theta_c, theta_s, max_hardening, hardening_factor, enable=True):
        """Clamped plasticity with hardening after every G2P (MPM_PLASTICITY_CLAMP)."""
        _check(lib().mpmx_set_plasticity(self._h, 1 if enable else 0, float(theta_c), float(theta_s),
                                         float(max_hardening), float(hardening_factor)))

    def get_plastic_state(self):
        """dict Fp (n, dim*dim), mu, lambda of the particles (user order)."""
        d, n = self.dim, self.n
        Fp = np.zeros((n, d * d), self.dtype)
        mu = np.zeros(n, self.dtype)
        la = np.zeros(n, self.dtype)
        _check(lib().mpmx_get_plastic_state(self._h, _p(Fp), _p(mu), _p(la)))
        return {"Fp": Fp, "mu": mu, "lambda": la}

    def set_drucker_prager(self, a0, a1=0.0, a3=0.0, a4=0.0, cohesion=0.0, volume_correction=True, enable=True):
        """Drucker-Prager plasticity after every G2P (MPM_DRUCKER_PRAGER); friction angle in degrees."""
        _check(lib().mpmx_set_drucker_prager(self._h, 1 if enable else 0, float(a0), float(a1), float(a3), float(a4),
                                             float(cohesion), 1 if volume_correction else 0))

    def get_drucker_prager_state(self):
        """dict plastic_def, rho_F, cohesion, vc of the particles (user order)."""
        out = {k: np.zeros(self.n, self.dtype) for k in ("plastic_def", "rho_F", "cohesion", "vc")}
        _check(lib().mpmx_get_drucker_prager_state(self._h, *(_p(out[k]) for k in ("plastic_def", "rho_F", "cohesion",
                                                                                  "vc"))))
        return out

    def tape_begin(self, checkpoint_every=20):
        _check(lib().mpmx_tape_begin(self._h, int(checkpoint_every)))

    def backward(self, dX=None, dV=None, dF=None, dB=None):
        """Reverse pass over the recorded steps.  Inputs are adjoints of the
        final state (None = zero); returns dict X, V, F, B (adjoints of the
        state at tape_begin) and mu, lambda (float64)."""
        d, n = self.dim, self.n
        ins = [None if a is None else self._arr(a, (n, c)) for a, c in ((dX, d), (dV, d), (dF, d * d), (dB, d * d))]
        outs = [np.zeros((n, c), self.dtype) for c in (d, d, d * d, d * d)]
        gmu = np.zeros(n)
        gla = np.zeros(n)
        _check(lib().mpmx_backward(self._h, *[_p(a) for a in ins], *[_p(a) for a in outs], _p(gmu), _p(gla)))
        return dict(X=outs[0], V=outs[1], F=outs[2], B=outs[3], mu=gmu, **{"lambda": gla})

    def enable_timing(self, on=True):
        """on: True/1 every phase, 2 the fused tile kernel only, False/0 off."""
        _check(lib().mpmx_enable_timing(self._h, int(on)))

    def reset_timing(self):
        _check(lib().mpmx_reset_timing(self._h))

    def phase_time(self, phase):
        t = np.zeros(1)
        c = np.zeros(1, np.int64)
        _check(lib().mpmx_phase_time(self._h, phase.encode(), _p(t), _p(c)))
        return float(t[0]), int(c[0])

    def stress(self, F, volume, mu, lam):
        d = self.dim
        F = self._arr(F, (-1, d * d))
        n = F.shape[0]
        vol, mu, lam = (self._arr(np.broadcast_to(x, (n,)), (n,)) for x in (volume, mu, lam))
        tau = np.zeros_like(F)
        _check(lib().mpmx_stress(self._h, n, _p(F), _p(vol), _p(mu), _p(lam), _p(tau)))
        return tau

    def svd(self, F):
        d = self.dim
        F = self._arr(F, (-1, d * d))
        n = F.shape[0]
        U = np.zeros_like(F)
        V = np.zeros_like(F)
        S = np.zeros((n, d), self.dtype)
        _check(lib().mpmx_svd(self._h, n, _p(F), _p(U), _p(S), _p(V)))
        return U, S, V

    def sort_keys(self, keys, key_bits):
        keys = np.ascontiguousarray(keys, np.uint32)
        n = keys.shape[0]
        ko = np.zeros(n, np.uint32)
        oo = np.zeros(n, np.uint32)
        _check(lib().mpmx_sort_keys(self._h, n, _p(keys), key_bits, _p(ko), _p(oo)))
        return ko, oo

    def sort_keys_noalloc(self, n, keys_dev, key_bits, keys_out_dev, order_out_dev):
        """mpmx_sort_keys_noalloc on device addresses (ints); queued, not waited for."""
        _check(lib().mpmx_sort_keys_noalloc(self._h, int(n), ctypes.c_void_p(keys_dev), int(key_bits),
                                            ctypes.c_void_p(keys_out_dev), ctypes.c_void_p(order_out_dev)))


def seed_particles(dim, cells, lo, hi, shapes, ppc_axis=2, seed=1234, precision=FP32, jitter=True, poisson=False,
                   particles_per_cell=0.0):
    """Jittered seeding of shapes (Poisson-disk seeding with ``poisson``), see include/mpmx_scene.h.

    ``shapes``: list of dicts {type: SHAPE_BOX|SHAPE_SPHERE, a: 3 floats,
    b: 3 floats (box max or [radius]), density, E, nu, v: velocity}.
    Returns a dict of particle arrays.
    """
    cells = [cells] * dim if np.isscalar(cells) else list(cells)
    ns = len(shapes)
    st = np.array([s["type"] for s in shapes], np.int32)
    sa = np.array([list(s["a"]) + [0.0] * (3 - len(s["a"])) for s in shapes], np.float64)
    sb = np.array([list(s["b"]) + [0.0] * (3 - len(s["b"])) for s in shapes], np.float64)
    den = np.array([s.get("density", 1000.0) for s in shapes], np.float64)
    E = np.array([s["E"] for s in shapes], np.float64)
    nu = np.array([s["nu"] for s in shapes], np.float64)
    vel = np.array([list(s.get("v", [0.0] * 3)) + [0.0] * (3 - len(s.get("v", [0.0] * 3))) for s in shapes],
                   np.float64)
    d = SeedDesc()
    d.dim = dim
    for a in range(3):
        d.cells[a] = cells[a] if a < dim else 1
        d.lo[a] = lo[a] if a < dim else 0.0
        d.hi[a] = hi[a] if a < dim else 1.0
    d.ppc_axis, d.seed, d.jitter, d.nshapes = ppc_axis, seed, 1 if jitter else 0, ns
    keep = (st, sa, sb, den, E, nu, vel)
    d.shape_type, d.shape_a, d.shape_b, d.density, d.youngs, d.poisson, d.velocity = [
        a.ctypes.data for a in keep]
    # Upper bound on the count (cells overlapping each shape's bounding box),
    # so that one pass of the generator fills the arrays.
    h = (np.asarray(hi[:dim], float) - np.asarray(lo[:dim], float)) / np.asarray(cells[:dim], float)
    cap = 0
    for s in shapes:
        a = np.asarray(s["a"][:dim], float)
        ext = np.asarray(s["b"][:dim], float) - a if s["type"] == SHAPE_BOX else np.full(dim, 2 * s["b"][0])
        cap += int(np.prod(np.ceil(ext / h) + 2)) * ppc_axis ** dim
    if poisson:  # Seed_Particles_Poisson: count first, then fill
        cap = lib().mpmx_seed_particles_poisson(ctypes.byref(d), precision, float(particles_per_cell), 0,
                                                *([None] * 9))
        if cap < 0:
            raise MpmxError(-1, "bad seeding description")
    dt = np.float32 if precision == FP32 else np.float64
    out = dict(X=np.zeros((cap, dim), dt), V=np.zeros((cap, dim), dt), F=np.zeros((cap, dim * dim), dt),
               B=np.zeros((cap, dim * dim), dt), mass=np.zeros(cap, dt), volume=np.zeros(cap, dt),
               mu=np.zeros(cap, dt), lam=np.zeros(cap, dt), valid=np.zeros(cap, np.uint8))
    ptrs = [_p(out[k]) for k in ("X", "V", "F", "B", "mass", "volume", "mu", "lam", "valid")]
    if poisson:
        n = lib().mpmx_seed_particles_poisson(ctypes.byref(d), precision, float(particles_per_cell), cap, *ptrs)
    else:
        n = lib().mpmx_seed_particles(ctypes.byref(d), precision, cap, *ptrs)
    del keep
    if n < 0:
        raise MpmxError(-1, "bad seeding description")
    if n > cap:
        raise MpmxError(-5, "seeding bound too small")
    return {k: v[:n] for k, v in out.items()}


def random_boxes(seed, nboxes=12, side_min=0.08, side_max=0.2, region_lo=0.1, region_hi=0.9,
                 E_min=1e3, E_max=1e5, nu_min=0.2, nu_max=0.4, max_tries=10000, cubes=True):
    lo = np.zeros((nboxes, 3))
    hi = np.zeros((nboxes, 3))
    E = np.zeros(nboxes)
    nu = np.zeros(nboxes)
    k = lib().mpmx_random_boxes(seed, nboxes, side_min, side_max, region_lo, region_hi, E_min, E_max, nu_min,
                                nu_max, max_tries, 1 if cubes else 0, _p(lo), _p(hi), _p(E), _p(nu))
    if k < 0:
        raise MpmxError(-1, "bad box description")
    return lo[:k], hi[:k], E[:k], nu[:k]
