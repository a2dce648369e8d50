// PhysBAM particle (PARTICLES<TV>::Write/Read version 2) and restart files.
#pragma once
#include <cstdint>

#include "mpmx.h"

namespace mpmx {
namespace io {
// Writes n particles' named arrays (any order; stored sorted by name).
void write_particles(const char* path, int dim, int precision, int64_t n, const mpmx_particle_array* arrays, int count);
// Returns the particle count; fills the named arrays that are present (found[i]
// = 1; a missing one is an error when found is null), converting the file's
// scalar type to `precision`.  count = 0 only reads the header.
int64_t read_particles(const char* path, int dim, int precision, mpmx_particle_array* arrays, int count, int* found);
void write_restart(const char* path, int precision, double time);
double read_restart(const char* path);
}  // namespace io
}  // namespace mpmx
