/*
 * mpmx — MI355X (gfx950) explicit APIC MPM time step behind a C ABI.
 *
 * One context holds one MPM scene on one GPU: the grid description
 * (GRID<TV> with ghost 3), the particle state (MPM_PARTICLES<TV>), the force
 * set (one MPM_FINITE_ELEMENTS with per-particle mu/lambda plus MPM_GRAVITY)
 * and the collision objects (axis-aligned boxes), and advances it with the
 * symplectic-Euler step of MPM_DRIVER<TV>::Advance_One_Time_Step.
 *
 * Reference paths below are relative to PhysBAM's Public_Library/.
 *
 * Conventions
 *   - every call returns MPMX_OK (0) or a negative MPMX_E* code; the message
 *     of the last failure on the calling thread is mpmx_last_error();
 *   - host arrays are caller-owned and copied in/out; device buffers are owned
 *     by the context;
 *   - floating point arrays have the context precision (float for
 *     MPMX_FP32, double for MPMX_FP64), vectors are d contiguous values per
 *     particle and d x d matrices are column-major per particle
 *     (Core/Matrices/MATRIX_3X3.h:115-119), exactly as MPM_PARTICLES stores X,
 *     V, F and B;
 *   - a context is externally synchronised and runs on one HIP stream; calls
 *     return after their GPU work has finished unless named *_async.
 */
#ifndef MPMX_H
#define MPMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPMX_OK 0
#define MPMX_EINVAL (-1)   /* bad argument */
#define MPMX_EHIP (-2)     /* HIP runtime failure */
#define MPMX_ESTATE (-3)   /* call not valid in the current state */
#define MPMX_ERANGE (-4)   /* particle outside the ghosted grid */
#define MPMX_EINTERNAL (-5)

#define MPMX_FP32 4
#define MPMX_FP64 8

/* Constitutive models (Deformables/Constitutive_Models). */
#define MPMX_FIXED_COROTATED 0 /* COROTATED_FIXED.cpp:59-65 */
#define MPMX_NEO_HOOKEAN 1     /* NEO_HOOKEAN.cpp:44-81 */
#define MPMX_HENCKY 2          /* ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.cpp:60-70 */

/* Collision types, MPM_COLLISION_OBJECT<TV>::COLLISION_TYPE
   (Hybrid_Methods/Collisions/MPM_COLLISION_OBJECT.h:19). */
#define MPMX_STICK 0
#define MPMX_SLIP 1
#define MPMX_SEPARATE 2

typedef struct mpmx_ctx mpmx_ctx;

/* GRID<TV>::Initialize(counts, RANGE(lo, hi), MAC=true) with ghost 3
   (Grid_Tools/Grids/GRID.cpp:13-27; MPM_EXAMPLE.h:86). */
typedef struct mpmx_grid_desc {
  int dim;         /* 2 or 3 */
  int precision;   /* MPMX_FP32 or MPMX_FP64 */
  int cells[3];    /* cells per axis (unused axes ignored) */
  double lo[3];    /* domain min corner */
  double hi[3];    /* domain max corner */
} mpmx_grid_desc;

/* Message of the last failed call on this thread ("" if none). */
const char* mpmx_last_error(void);

/* Create a context on the current HIP device.  Replaces the construction of
   MPM_EXAMPLE<TV> + MPM_DRIVER<TV> (Examples_And_Drivers/MPM_EXAMPLE.cpp:21-31,
   MPM_DRIVER.cpp:47-55) and the grid set-up of MPM_DRIVER::Initialize
   (MPM_DRIVER.cpp:102-108). */
int mpmx_create(const mpmx_grid_desc* grid, mpmx_ctx** out);
int mpmx_destroy(mpmx_ctx* ctx);

/* Batch of independent scenes in one context (no counterpart in the reference,
   which runs one MPM_EXAMPLE per process): scene s has its own grid grids[s]
   (same dim and precision for all) and its own particles; every step advances
   all scenes with the same launches.  Forces, gravity, collision boxes and dt
   are shared and act in each scene's own coordinates.  Particles are set with
   mpmx_set_batch_particles (arrays concatenated in scene order, counts[s]
   particles for scene s) and read back by mpmx_get_particles in that order.
   mpmx_get_grid is for single-scene contexts. */
int mpmx_create_batch(const mpmx_grid_desc* grids, int nscenes, mpmx_ctx** out);
int mpmx_scene_count(const mpmx_ctx* ctx);
int mpmx_set_batch_particles(mpmx_ctx* ctx, int nscenes, const int64_t* counts, const void* X, const void* V,
                             const void* F, const void* B, const void* mass, const void* volume, const void* mu,
                             const void* lambda, const uint8_t* valid);

/* Particle state in, MPM_PARTICLES<TV> layout (Examples_And_Drivers/MPM_PARTICLES.h:17-44):
   X, V: n*d; F, B: n*d*d (column-major); mass, volume, mu, lambda: n; valid: n
   bytes (nonzero = simulated).  mu/lambda are the per-particle Lame parameters
   bound by MPM_FINITE_ELEMENTS (Forces/MPM_FINITE_ELEMENTS.cpp:30-31).
   Replaces STANDARD_TESTS_BASE::Add_Particle + Set_Lame_On_Particles
   (Projects/mpm/STANDARD_TESTS_BASE.cpp:285-310, :567-586).  Fails with
   MPMX_ERANGE if a valid particle's stencil would leave the ghosted grid. */
int mpmx_set_particles(mpmx_ctx* ctx, int64_t n, const void* X, const void* V, const void* F, const void* B,
                       const void* mass, const void* volume, const void* mu, const void* lambda,
                       const uint8_t* valid);

/* Particle state out (any pointer may be NULL), in the order given to
   mpmx_set_particles. */
int mpmx_get_particles(mpmx_ctx* ctx, void* X, void* V, void* F, void* B, uint8_t* valid);
int64_t mpmx_particle_count(const mpmx_ctx* ctx);
/* mass, volume and the Lame parameters mu, lambda of the particles (any
   pointer may be NULL), in the same order; with clamped plasticity mu and
   lambda are the hardened values. */
int mpmx_get_material(mpmx_ctx* ctx, void* mass, void* volume, void* mu, void* lambda);

/* Elastic force: one MPM_FINITE_ELEMENTS with the given model over all
   particles (STANDARD_TESTS_BASE::Add_Fixed_Corotated / Add_Neo_Hookean,
   STANDARD_TESTS_BASE.cpp:372-392; MPMX_HENCKY is the St. Venant-Kirchhoff
   Hencky-strain model that Add_Drucker_Prager uses, :409-432).  neo_threshold is NEO_HOOKEAN's
   failure_threshold (default 0.25, NEO_HOOKEAN.h:33). */
int mpmx_set_constitutive(mpmx_ctx* ctx, int model, double neo_threshold);

/* MPM_GRAVITY (Forces/MPM_GRAVITY.cpp:62-69); g has dim entries. */
int mpmx_set_gravity(mpmx_ctx* ctx, const double* g);

/* MPM_EXAMPLE::Add_Collision_Object with an ANALYTIC_IMPLICIT_OBJECT<RANGE<TV>>
   (MPM_EXAMPLE.cpp:331-338; MPM_EXAMPLE.h:153-156); box corners have dim
   entries.  At most 16 objects. */
int mpmx_add_collision_box(mpmx_ctx* ctx, int type, double friction, const double* lo, const double* hi);
int mpmx_clear_collision_objects(mpmx_ctx* ctx);

/* nsteps calls of MPM_DRIVER::Advance_One_Time_Step with example.dt = dt
   (MPM_DRIVER.cpp:141-157), advancing example.time by dt each. */
int mpmx_step(mpmx_ctx* ctx, double dt, int nsteps);
/* Same, enqueued on the context stream without waiting. */
int mpmx_step_async(mpmx_ctx* ctx, double dt, int nsteps);
int mpmx_synchronize(mpmx_ctx* ctx);

/* The step one named step at a time, for callers that fire callbacks around
   each (MPM_EXAMPLE::Add_Callbacks, MPM_EXAMPLE.cpp:350-356, fired by
   MPM_DRIVER::Step, MPM_DRIVER.cpp:1380-1393).  The stages are the named
   steps of Advance_One_Time_Step in order (MPM_DRIVER.cpp:141-157):
     0 "simulated-particles"  (nothing to do: invalid particles are skipped)
     1 "update-weights"       (nothing to do: the kernels weight from X)
     2 "prepare-scatter"      (the bins, rebuilt when due, and the grid tiles)
     3 "p2g"
     4 "forces"               (grid update: gravity, collision objects, friction)
     5 "reduce-dt"            (the strong and sound-speed CFL when set)
     6 "g2p"
     7 "plasticity-hardening" (then the step ends and the time advances)
   Stage 0 opens a step with example.dt = dt (dt is ignored by the others);
   each call must take the next stage (MPMX_ESTATE otherwise, and mpmx_step
   fails while a step is open) and returns after its GPU work.  Such a step
   runs the unfused kernels.  Between stages 3 and 4, mpmx_get_grid returns
   P2G's velocities (momentum / mass, before any force), as example.velocity
   holds them there.
   mpmx_stage_name returns the stage's name, or NULL. */
#define MPMX_STAGES 8
int mpmx_step_stage(mpmx_ctx* ctx, double dt, int stage);
const char* mpmx_stage_name(int stage);

/* Simulation time (example.time). */
double mpmx_time(const mpmx_ctx* ctx);

/* FP64 grid totals after the last step's grid update: mass and momentum
   sum_i m_i v_i over nodes with nonzero mass, i.e. MPM_EXAMPLE::
   Total_Grid_Linear_Momentum of the velocity used by G2P (MPM_EXAMPLE.cpp:
   226-241).  momentum has dim entries.  Computed only when enabled. */
int mpmx_enable_grid_stats(mpmx_ctx* ctx, int enable);
int mpmx_grid_stats(mpmx_ctx* ctx, double* mass, double* momentum);

/* FP64 particle totals over valid particles: mass and sum m_p v_p
   (MPM_EXAMPLE::Total_Particle_Linear_Momentum, MPM_EXAMPLE.cpp:209-225). */
int mpmx_particle_stats(mpmx_ctx* ctx, double* mass, double* momentum, int64_t* valid_count);

/* Conservation totals in FP64, the quantities -print_stats reports around
   each step (MPM_DRIVER::Print_Particle_Stats, Print_Grid_Stats and
   Print_Energy_Stats, Hybrid_Methods/Examples_And_Drivers/MPM_DRIVER.cpp:
   1159-1213).  Twelve doubles, in this order in memory: */
typedef struct mpmx_totals {
  double mass;       /* sum of particle (node) masses */
  double linear[3];  /* linear momentum (dim entries) */
  double angular[3]; /* angular momentum about the origin (2D: angular[0]) */
  double kinetic;    /* kinetic energy */
  double potential;  /* particles: elastic + gravity potential energy; grid: 0 */
  double elastic;    /* particles: sum volume * Energy_Density(singular values of F) */
  double gravity;    /* particles: -sum m x . g */
  double count;      /* valid particles, or nodes with nonzero mass */
} mpmx_totals;

/* Particle totals of the current state over valid particles:
   Total_Particle_Linear_Momentum, Total_Particle_Angular_Momentum (sum
   m x cross v - m B:eps, eps the permutation tensor) and
   Total_Particle_Kinetic_Energy (sum m/2 |v|^2 + m/2 (B D^-1):B with
   D^-1 = 4/dx^2 of the quadratic spline) (MPM_EXAMPLE.cpp:209-225, :248-263,
   :302-315), and MPM_EXAMPLE::Potential_Energy (MPM_EXAMPLE.cpp:123-133) of
   the elastic force (MPM_FINITE_ELEMENTS.cpp:72-80) and gravity
   (MPM_GRAVITY.cpp:47-55).  Deterministic for a given particle storage order. */
int mpmx_particle_totals(mpmx_ctx* ctx, mpmx_totals* out);

/* Grid totals of the last step, recorded by every step while grid stats are
   enabled (mpmx_enable_grid_stats): which = 0 after Particle_To_Grid (node
   velocities p/m before any force, with the reflection boundary conditions'
   mass and momentum reflection, as example.velocity holds them after the
   "p2g" step), which = 1 after the grid update (the velocity G2P reads):
   Total_Grid_Linear_Momentum, Total_Grid_Angular_Momentum and
   Total_Grid_Kinetic_Energy (MPM_EXAMPLE.cpp:226-241, :268-297) over the
   nodes with nonzero mass, node locations the MAC cell centres.  While grid
   stats are enabled every step takes the unfused kernels and one more P2G
   (without the elastic impulse) for the which = 0 totals. */
int mpmx_grid_totals(mpmx_ctx* ctx, int which, mpmx_totals* out);

/* Dense grid arrays of the last step, indexed like MPM_EXAMPLE::mass and
   ::velocity over Domain_Indices(ghost=3): (cells+6)^dim nodes, row-major
   with the last axis fastest (Core/Arrays_Nd/ARRAYS_ND_BASE.h:88-93).
   mass: nodes values; velocity: nodes*dim (the velocity G2P reads).  Nodes
   with zero mass read as zero. */
int64_t mpmx_grid_node_count(const mpmx_ctx* ctx);
int mpmx_get_grid(mpmx_ctx* ctx, void* mass, void* velocity);

/* Binning (replaces GATHER_SCATTER::Prepare_Scatter, Hybrid_Methods/
   Iterators/GATHER_SCATTER.cpp:32-86).  Particles are binned by the grid tile
   holding their stencil base; a particle that leaves its binned tile (a
   "mover") is transferred on a slower path until the bins are rebuilt.  The
   bins are rebuilt when the mover count two steps earlier exceeds
   mover_fraction * n, or after max_interval steps (0 = no limit).  Defaults:
   1/128 and 0.  Results do not depend on the policy beyond floating-point
   summation order. */
int mpmx_set_rebin_policy(mpmx_ctx* ctx, double mover_fraction, int max_interval);
/* Bin rebuilds so far, and the movers and FP64-SVD stress particles of the
   last step (any pointer may be NULL). */
int mpmx_bin_stats(mpmx_ctx* ctx, int64_t* rebins, int64_t* movers, int64_t* slow);

/* Tuning: at a bin rebuild, reorder particle storage by tile if at least
   `interval` steps (>=1) have passed since the last reorder. */
int mpmx_set_reorder_interval(mpmx_ctx* ctx, int interval);

/* Launch configuration of the fused step (no counterpart in the reference,
   whose step runs on the host).  graph = 1: the fused steps of mpmx_step and
   mpmx_step_async run four at a time from one captured hipGraph, re-captured
   when the arguments change (used while timing and tapes are off; other
   steps run as before); bins are rebuilt between groups of four when due.
   tile_workgroups: workgroups the tile kernels launch, each striding over the
   tile list (0 = one per tile up to 8192; the resident count of the fused
   kernel or a small multiple of it makes them persistent).  tile_threads:
   workgroup size of the fused tile kernel, 128 or 256 (0 = 128): the tile's
   particles are taken that many at a time, and each cell's particles of a
   batch are added in rounds of tile_threads / 64.  Results do not depend on
   the configuration beyond floating-point summation order.  For a batch
   context all scenes run in these same launches. */
typedef struct mpmx_launch_config {
  int graph;
  int tile_workgroups;
  int tile_threads;
} mpmx_launch_config;
int mpmx_set_launch_config(mpmx_ctx* ctx, const mpmx_launch_config* cfg);
int mpmx_get_launch_config(mpmx_ctx* ctx, mpmx_launch_config* cfg);

/* Autotuner of the launch configuration (no counterpart in the reference):
   times trial_steps (>= 8) fused steps of dt of the context's own particles
   after 4 untimed ones, for each candidate in this order: graph off, then on;
   within each, tile_threads 128, then 256; within each, tile_workgroups 0,
   then 1, 2 and 4 times the fused kernel's resident workgroup count (16
   candidates); applies the fastest by wall clock, returns it in
   *chosen, and writes each candidate's milliseconds per step to times_ms
   (up to max_times entries; *ntimes = number of candidates).  The particles,
   the simulation time and the timing switch are restored afterwards; the
   bins are rebuilt at the next step.  Needs particles and a step the fused
   kernels take (no plasticity, dt limits, reflection, grid stats or tape:
   MPMX_ESTATE). */
int mpmx_autotune(mpmx_ctx* ctx, double dt, int trial_steps, mpmx_launch_config* chosen, double* times_ms,
                  int max_times, int* ntimes);

/* Adaptive time step.
   mpmx_grid_v_upper_bound replaces MPM_DRIVER::Grid_V_Upper_Bound
   (Hybrid_Methods/Examples_And_Drivers/MPM_DRIVER.cpp:1125-1138, affine CFL
   on, the default of MPM_EXAMPLE.h:116): per scene, the max over valid
   particles of |v| + 6 sqrt(dim) min(1/dx) |B|_F, computed in the context's
   precision with the reference's rounding (bit-identical for the same
   particle state).  `bound` receives one double per scene.
   mpmx_compute_dt replaces MPM_DRIVER::Compute_Dt (MPM_DRIVER.cpp:1098-1104):
   per scene, cfl*min(dx)/bound when the bound exceeds cfl*min(dx)/max_dt,
   else max_dt.
   mpmx_advance_frames replaces the substep loop of
   MPM_DRIVER::Simulate_To_Frame (MPM_DRIVER.cpp:173-197) for `nframes`
   frames of length frame_dt from the context time: each substep takes
   Compute_Dt clamped to [min_dt, max_dt], a substep within 1e-4 dt of the
   frame end lands on it, and the one before the last splits the remainder in
   half (all in the context's precision).  The substep dts are written to
   `dts` (up to dts_capacity of them; NULL with capacity 0 skips them) and
   their number to *nsubsteps.  Single-scene contexts only (MPMX_ESTATE for
   batches); strong and sound-speed CFL (off by default in the reference) are
   not provided.  MPMX_ESTATE if a substep dt is not positive. */
int mpmx_grid_v_upper_bound(mpmx_ctx* ctx, double* bound);
int mpmx_compute_dt(mpmx_ctx* ctx, double cfl, double max_dt, double* dt);
int mpmx_advance_frames(mpmx_ctx* ctx, int nframes, double frame_dt, double cfl, double min_dt, double max_dt,
                        double* dts, int64_t dts_capacity, int64_t* nsubsteps);

/* Reflection boundary conditions on the domain sides instead of wall
   objects: example.reflection_bc with side_bc_type (MPM_EXAMPLE.h:106-109,
   :134), as -use_reflect makes Add_Walls set them (Projects/mpm/
   STANDARD_TESTS_BASE.cpp:487-505).  side_types holds 2*dim entries for
   -x, +x, -y, +y[, -z, +z]: MPMX_BC_FREE, MPMX_BC_SLIP, MPMX_BC_NOSLIP or
   MPMX_BC_SEPARATE.  full = 0: the P2G mass and momentum, the forces and the
   final velocities are reflected across the sides (Reflect_Boundary_Mass_
   Momentum, _Force, _Velocity, MPM_DRIVER.cpp:310-342, :395-446); full = 1
   (-use_reflect_friction): the velocities of node pairs across a side are
   mixed by mass after the collision objects, separating sides applying
   Coulomb friction `friction` (Reflect_Boundary_Friction, :343-393).
   Particles leaving through a side whose type is not free are mirrored back
   (Reflect_Or_Invalidate_Particle, :1219-1239).  Needs at least 6 cells per
   axis; not for batches, tapes or the dt limits, and the full variant not
   with collision objects that have friction.  Steps with it take the unfused
   kernels; the tiles along the sides stay in the grid every step. */
#define MPMX_BC_FREE 0
#define MPMX_BC_SLIP 1
#define MPMX_BC_NOSLIP 2
#define MPMX_BC_SEPARATE 3
int mpmx_set_reflection(mpmx_ctx* ctx, int enable, int full, double friction, const int* side_types);

/* Strong and sound-speed CFL: MPM_DRIVER::Reduce_Dt (MPM_DRIVER.cpp:519-545)
   between the grid update and G2P of every step, as -strong_cfl and
   -sound_cfl turn it on (Projects/mpm/STANDARD_TESTS_BASE.cpp:128-129).
   strong_cfl: Grid_To_Particle_Limit_Dt (:572-620) lowers dt so that no entry
   of dt grad v changes by more than cfl_F and no particle moves more than cfl
   (in the domain's length units, as example.cfl) along an axis; sound_cfl:
   Limit_Dt_Sound_Speed (:624-637) caps dt at cfl_sound min dx / c_max with
   the largest particle sound speed (ISOTROPIC_CONSTITUTIVE_MODEL::
   Sound_Speed).  The result is raised to min_dt; when it is below 0.99 dt
   the step's grid velocity changes are scaled to it and G2P takes it, and
   mpmx_step advances the time by it (mpmx_last_dt reads it back);
   mpmx_advance_frames then continues the frame as Simulate_To_Frame does
   (:190-194), using its own cfl and min_dt.  Defaults cfl 1, cfl_F 0.5,
   cfl_sound 0.9, min_dt 0 (MPM_EXAMPLE.h:95-97); both off.  Steps with a
   limit on read two values back from the GPU and take the unfused kernels;
   not for batches or tapes.  mpmx_max_sound_speed is Compute_Max_Sound_Speed
   (:893-918) of the current particles. */
int mpmx_set_dt_limits(mpmx_ctx* ctx, int strong_cfl, int sound_cfl, double cfl, double cfl_F, double cfl_sound,
                       double min_dt);
int mpmx_max_sound_speed(mpmx_ctx* ctx, double* out);
int mpmx_last_dt(mpmx_ctx* ctx, double* out);

/* Plasticity: MPM_PLASTICITY_CLAMP with hardening
   (Hybrid_Methods/Forces/MPM_PLASTICITY_CLAMP.cpp:31-45 and
   MPM_PLASTICITY_MODEL.cpp:34-48, as Projects/mpm/STANDARD_TESTS_BASE.cpp:
   465-486 Add_Clamped_Plasticity adds it to the explicit step).  After every
   G2P the singular values of each simulated particle's F are clamped to
   [1 - theta_c, 1 + theta_s]; when any changes, mu and lambda become mu0 and
   lambda0 times max(1e-8, exp(min(max_hardening, hardening_factor (1 - det
   Fp)))), F = U clamp(S) V^T and Fp = V clamp(S)^-1 U^T F Fp.  Enabling sets
   Fp = I and mu0, lambda0 = the current mu, lambda of every particle (and so
   does each later mpmx_set_particles); enable = 0 turns it off.  Steps with
   plasticity take the unfused G2P/P2G kernels.  mpmx_get_plastic_state
   returns Fp (dim*dim per particle, column-major), mu and lambda (NULL
   skips).  Not differentiable: mpmx_tape_begin rejects it. */
int mpmx_set_plasticity(mpmx_ctx* ctx, int enable, double theta_c, double theta_s, double max_hardening,
                        double hardening_factor);
int mpmx_get_plastic_state(mpmx_ctx* ctx, void* Fp, void* mu, void* lambda);

/* Plasticity: MPM_DRUCKER_PRAGER (sand; Hybrid_Methods/Forces/
   MPM_DRUCKER_PRAGER.cpp:40-131, as STANDARD_TESTS_BASE.cpp:407-431
   Add_Drucker_Prager adds it).  After every G2P the log singular values of
   each simulated particle's F are returned onto the Drucker-Prager cone whose
   friction angle a0 + (a1 q - a4) exp(-a3 q) degrees hardens with the
   accumulated plastic deformation q; `cohesion` is sigma_Y, and
   volume_correction keeps the volume lost in tension (the reference's vc,
   on unless -no_vc).  Used with MPMX_HENCKY (St. Venant-Kirchhoff with Hencky
   strain), the model the reference pairs it with.  Enabling sets Fp = I,
   q = 0, vc = 0 and sigma_Y = cohesion for every particle (and so does each
   later mpmx_set_particles); it replaces clamped plasticity, and enable = 0
   turns it off.  mpmx_get_plastic_state returns Fp; mpmx_get_drucker_prager_
   state returns q, rho_F (the cone's slope), sigma_Y and vc per particle
   (NULL skips).  Steps with it take the unfused kernels; not
   differentiable. */
int mpmx_set_drucker_prager(mpmx_ctx* ctx, int enable, double a0, double a1, double a3, double a4, double cohesion,
                            int volume_correction);
int mpmx_get_drucker_prager_state(mpmx_ctx* ctx, void* plastic_def, void* rho_F, void* cohesion, void* vc);

/* Particle and restart files in PhysBAM's own format, for restarts and for
   the tools that read PhysBAM output (the viewer, mpm2partio).
   mpmx_write_particles_file is PARTICLES<TV>::Write (Tools/Particles/
   PARTICLES.cpp:362-373) behind a FILE_OSTREAM (Core/Read_Write/
   TYPED_STREAM.cpp:30-36): a byte saying whether the stream holds doubles,
   version 2, the particle count and each named array (stored in name order)
   with the coded type id of ATTRIBUTE_INFO::Encode (:15-25).  It writes the
   `count` arrays given, each n elements in the precision given; a path ending
   in ".gz" is gzip-compressed (Safe_Open_Output, FILE_UTILITIES.cpp:237-261).
   mpmx_read_particles_file is PARTICLES<TV>::Read for version 2 (:323-358):
   it returns the particle count in *n (arrays = NULL, count = 0 reads only
   the header) and fills the named arrays found in the file, converting float
   and double; found[i] = 1 for each one present (found = NULL makes a missing
   array an error); other arrays are skipped. */
#define MPMX_ARRAY_SCALAR 0  /* one value per particle (ARRAY<T>) */
#define MPMX_ARRAY_VECTOR 1  /* dim values (ARRAY<TV>) */
#define MPMX_ARRAY_VECTOR3 2 /* 3 values (ARRAY<VECTOR<T,3> >, e.g. color) */
#define MPMX_ARRAY_MATRIX 3  /* dim*dim values, column-major (ARRAY<MATRIX<T,d> >) */
#define MPMX_ARRAY_BOOL 4    /* one byte (ARRAY<bool>) */
typedef struct mpmx_particle_array {
  const char* name; /* attribute name, e.g. "X", "mass", "valid" */
  int kind;         /* MPMX_ARRAY_* */
  void* data;       /* n elements in the call's precision */
} mpmx_particle_array;
int mpmx_write_particles_file(const char* path, int dim, int precision, int64_t n, const mpmx_particle_array* arrays,
                              int count);
int mpmx_read_particles_file(const char* path, int dim, int precision, int64_t* n, mpmx_particle_array* arrays,
                             int count, int* found);

/* restart_data: the time behind the same one-byte stream header, in the
   given precision (Write_To_File of example.time, MPM_EXAMPLE.cpp:66). */
int mpmx_write_restart_file(const char* path, int precision, double time);
int mpmx_read_restart_file(const char* path, double* time);

/* MPM_EXAMPLE::Write_Output_Files' particle and restart files for one frame
   directory (MPM_EXAMPLE.cpp:62-74, DEFORMABLE_BODY_COLLECTION.cpp:112):
   <frame_dir>/deformable_object_particles.gz with the arrays an MPM scene of
   the reference stores (X, V, F, B, color, mass, volume, valid, mu, lambda;
   with plasticity also Fp, mu0, lambda0 and, for Drucker-Prager,
   plastic_deformation, dp_rho_f, dp_cohesion, vcp), the link mpm_particles.gz
   to it, and <frame_dir>/restart_data.gz with the time.  color is written as
   zeros (rendering data this library does not keep).  mpmx_read_output_files
   is MPM_EXAMPLE::Read_Output_Files (:95-102): mpm_particles (or
   deformable_object_particles) and restart_data, gzipped or not, replace the
   particles (as mpmx_set_particles) and the time; with plasticity enabled the
   plastic arrays in the file are restored too. */
int mpmx_write_output_files(mpmx_ctx* ctx, const char* frame_dir);
int mpmx_read_output_files(mpmx_ctx* ctx, const char* frame_dir);

/* Differentiation (no counterpart in the reference, whose step has no
   adjoint).  mpmx_tape_begin records the current particle state and then a
   checkpoint every `checkpoint_every` steps of the following mpmx_step calls.
   mpmx_backward takes the adjoints of the final state (dL/dX, dL/dV, dL/dF,
   dL/dB in the mpmx_set_particles layout; NULL = zero), runs the reverse pass
   through every recorded step (recomputing each checkpoint segment forward;
   with two or more segments, on a second internal stream beside the reverse
   kernels of the segment after it), and returns the adjoints of the state at mpmx_tape_begin (NULL outputs are
   skipped) and dL/dmu, dL/dlambda per particle (double, accumulated in FP64).
   The particle state afterwards is the final state again; the tape is
   consumed.  Not supported: wall friction (rejected by mpmx_tape_begin) and
   the neo-Hookean clamp branch (MPMX_EINVAL from mpmx_backward).  The
   collision projection is differentiated with its face normal held fixed. */
int mpmx_tape_begin(mpmx_ctx* ctx, int checkpoint_every);
int mpmx_backward(mpmx_ctx* ctx, const void* dX, const void* dV, const void* dF, const void* dB, void* gX, void* gV,
                  void* gF, void* gB, double* gmu, double* glambda);

/* Per-phase GPU timing with HIP events on the context stream.  Phases:
   "rebin" (keys, sort, tile ranges; only on rebuild steps), "bin" (per-step
   grid-tile list and clearing), "p2g", "grid", "g2p", "reorder", "step",
   "adjoint" (a whole mpmx_backward pass), "dt" (the dt limits of a step),
   "g2p2g" (a fused step's G2P and next P2G with their mover side paths) and
   "g2p2g_kernel" (the fused kernel of that phase alone); "graph" counts the
   graph replays (groups of four fused steps) whether timing is on or not.
   mpmx_phase_time returns the accumulated milliseconds and launch count since
   the last reset.  enable: 0 off, 1 every phase, 2 only "g2p2g_kernel" (two
   events per fused step: the other phases' events cost stream time of their
   own); other values are MPMX_EINVAL. */
int mpmx_enable_timing(mpmx_ctx* ctx, int enable);
int mpmx_reset_timing(mpmx_ctx* ctx);
int mpmx_phase_time(mpmx_ctx* ctx, const char* phase, double* total_ms, int64_t* count);

/* Stand-alone stress kernel: tau_p = volume_p * P(F_p) F_p^T for n particles
   (the per-particle term of MPM_FINITE_ELEMENTS::Precompute + Add_Forces,
   Forces/MPM_FINITE_ELEMENTS.cpp:59-65, :91-96), host arrays in and out in
   the context precision.  F, tau: n*d*d column-major. */
int mpmx_stress(mpmx_ctx* ctx, int64_t n, const void* F, const void* volume, const void* mu,
                const void* lambda, void* tau);

/* Stand-alone SVD kernel F = U diag(s) V^T for n matrices, the convention of
   MATRIX<T,3>::Singular_Value_Decomposition (Core/Matrices/MATRIX_3X3.cpp:
   198-236): U, V rotations, |s0|>=|s1|>=|s2|, only the last may be negative. */
int mpmx_svd(mpmx_ctx* ctx, int64_t n, const void* F, void* U, void* s, void* V);

/* Stand-alone binning sort: stable sort of n uint32 keys (values = positions
   0..n-1) on the GPU; keys_out / order_out are host arrays of n entries.
   Only the low key_bits bits take part, and 0xffffffff sorts last. */
int mpmx_sort_keys(mpmx_ctx* ctx, int64_t n, const uint32_t* keys, int key_bits, uint32_t* keys_out,
                   uint32_t* order_out);

/* The binning sort without allocation.  No counterpart in the reference: it is
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  keys,
   keys_out and order_out are DEVICE arrays of n entries on the context's GPU;
   the sort is queued on the context stream and the call returns without
   waiting (mpmx_synchronize waits).  Results equal mpmx_sort_keys's bit for
   bit.  Its scratch (ping-pong arrays, histograms) is one static device array
   per GPU of about 8.5 MiB, kept until the process ends, so n is at most
   MPMX_SORT_NOALLOC_MAX_KEYS (MPMX_EINVAL above it, before any GPU work);
   every call on a GPU shares that array, so calls that overlap on one GPU
   (two contexts' streams) race.  It calls no hipMalloc or hipFree. */
#define MPMX_SORT_NOALLOC_MAX_KEYS (1 << 20)
int mpmx_sort_keys_noalloc(mpmx_ctx* ctx, int64_t n, const uint32_t* keys, int key_bits, uint32_t* keys_out,
                           uint32_t* order_out);

#ifdef __cplusplus
}
#endif

#endif /* MPMX_H */
