/*
 * Synthetic MPM scenes (host side, no GPU): jittered particle seeding and the
 * random box placement of the 12-box benchmark scene.
 *
 * This replaces the scene set-up of PhysBAM's Projects/mpm STANDARD_TESTS
 * (Seed_Particles + Add_Particle + Set_Lame_On_Particles,
 * Projects/mpm/STANDARD_TESTS_BASE.cpp:252-310, :567-586) for the benchmark
 * workloads.  All randomness is std::mt19937 with u = mt() / 2^32, the
 * generator and mapping of Core/Random_Numbers/RANDOM_NUMBERS.h:34-43.
 */
#ifndef MPMX_SCENE_H
#define MPMX_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPMX_SHAPE_BOX 0
#define MPMX_SHAPE_SPHERE 1

typedef struct mpmx_seed_desc {
  int dim;
  int cells[3];
  double lo[3], hi[3];
  int ppc_axis;          /* samples per cell per axis: 2 gives 4 (2D) / 8 (3D) per cell */
  uint32_t seed;
  int jitter;            /* 1: uniform in each sub-cell; 0: sub-cell centre */
  int nshapes;
  const int* shape_type;     /* MPMX_SHAPE_BOX | MPMX_SHAPE_SPHERE */
  const double* shape_a;     /* box min corner | sphere centre (3 per shape) */
  const double* shape_b;     /* box max corner | sphere radius in [0] (3 per shape) */
  const double* density;     /* per shape */
  const double* youngs;      /* per shape */
  const double* poisson;     /* per shape */
  const double* velocity;    /* per shape, 3 per shape */
} mpmx_seed_desc;

/* Jittered seeding.  Cells are visited row-major (last axis fastest); each is
   split into ppc_axis^dim sub-cells (row-major), one sample per sub-cell drawn
   with coordinates in axis order; a sample is kept if it lies in a shape (the
   first containing shape gives its material).  Samples outside the union
   bounding box of the shapes are drawn and skipped.  volume = h^d / ppc,
   mass = density * volume, F = I, B = 0, mu/lambda from (E, nu).
   Returns the particle count.  Arrays (precision 4 = float, 8 = double) are
   filled only if X is non-null and capacity >= count. */
int64_t mpmx_seed_particles(const mpmx_seed_desc* desc, int precision, int64_t capacity, void* X, void* V,
                            void* F, void* B, void* mass, void* volume, void* mu, void* lambda, uint8_t* valid);

/* Poisson-disk seeding, STANDARD_TESTS_BASE::Seed_Particles_Poisson
   (Projects/mpm/STANDARD_TESTS_BASE.cpp:223-234), the reference's default
   seeding: one RANDOM_NUMBERS stream seeded with desc->seed, a POISSON_DISK
   whose minimum distance comes from Set_Distance_By_Volume(dX.Product() /
   particles_per_cell) of the MAC grid (Geometry/Seeding/POISSON_DISK.cpp:
   130-137; sqrt(2/3 v) in 2D, cbrt(13/18 v) in 3D), and one Sample per shape
   in order over the shape's box, points outside a sphere then dropped (the
   last point swapped into the hole, as Remove_Index_Lazy).  All of it in the
   given precision, so that a seed gives the reference's points bit for bit.
   volume = dX.Product() / particles_per_cell and mass = density * volume in
   that precision; V, F = I, B = 0, mu, lambda as mpmx_seed_particles.
   particles_per_cell <= 0 takes ppc_axis^dim.  desc->jitter is not used.
   Returns the particle count (or -1); arrays are filled only if X is
   non-null and capacity >= count. */
int64_t mpmx_seed_particles_poisson(const mpmx_seed_desc* desc, int precision, double particles_per_cell,
                                    int64_t capacity, void* X, void* V, void* F, void* B, void* mass, void* volume,
                                    void* mu, void* lambda, uint8_t* valid);

/* Random non-overlapping axis-aligned boxes in [region_lo, region_hi]^3 with
   side lengths uniform in [side_min, side_max] (one draw per box if cubes is
   nonzero, else one per axis), placed by rejection (at most max_tries per
   box); Young's modulus log-uniform in [E_min, E_max], Poisson ratio uniform
   in [nu_min, nu_max].  Draw order per attempt: the side(s), then the three
   min-corner coordinates; per accepted box: E then nu.  Returns the number of
   boxes placed. */
int mpmx_random_boxes(uint32_t seed, int nboxes, double side_min, double side_max, double region_lo,
                      double region_hi, double E_min, double E_max, double nu_min, double nu_max, int max_tries,
                      int cubes, double* box_lo, double* box_hi, double* E, double* nu);

#ifdef __cplusplus
}
#endif

#endif /* MPMX_SCENE_H */
