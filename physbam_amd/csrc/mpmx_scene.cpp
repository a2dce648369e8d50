// Synthetic scene generation (host).  See include/mpmx_scene.h.
#include "mpmx_scene.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace {

struct Rng {
  std::mt19937 mt;
  explicit Rng(uint32_t seed) : mt(seed) {}
  double u() { return (double)mt() * (1.0 / 4294967296.0); }  // RANDOM_NUMBERS::Get_Number
  double uniform(double a, double b) { return a + (b - a) * u(); }
};

bool inside(const mpmx_seed_desc& d, int s, const double* x) {
  const double* a = d.shape_a + 3 * s;
  const double* b = d.shape_b + 3 * s;
  if (d.shape_type[s] == MPMX_SHAPE_BOX) {
    for (int k = 0; k < d.dim; k++)
      if (x[k] < a[k] || x[k] > b[k]) return false;
    return true;
  }
  double r2 = 0;
  for (int k = 0; k < d.dim; k++) r2 += (x[k] - a[k]) * (x[k] - a[k]);
  return r2 <= b[0] * b[0];
}

struct Sample {
  double x[3];
  int shape;
};

// One pass over the cells; returns the kept samples in seeding order.
std::vector<Sample> seed_samples(const mpmx_seed_desc& d) {
  const int dim = d.dim, q = d.ppc_axis;
  double h[3] = {1, 1, 1};
  for (int a = 0; a < dim; a++) h[a] = (d.hi[a] - d.lo[a]) / d.cells[a];
  int ppc = 1;
  for (int a = 0; a < dim; a++) ppc *= q;
  // sub-cell offsets, row-major (last axis fastest)
  std::vector<int> subs((size_t)ppc * 3, 0);
  for (int sub = 0; sub < ppc; sub++) {
    int r = sub;
    for (int a = dim - 1; a >= 0; a--) {
      subs[(size_t)sub * 3 + a] = r % q;
      r /= q;
    }
  }
  // union bounding box of the shapes: samples outside it cannot be kept
  double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
  for (int s = 0; s < d.nshapes; s++)
    for (int a = 0; a < dim; a++) {
      const double* sa = d.shape_a + 3 * s;
      const double* sb = d.shape_b + 3 * s;
      const double l = d.shape_type[s] == MPMX_SHAPE_BOX ? sa[a] : sa[a] - sb[0];
      const double u = d.shape_type[s] == MPMX_SHAPE_BOX ? sb[a] : sa[a] + sb[0];
      blo[a] = std::min(blo[a], l);
      bhi[a] = std::max(bhi[a], u);
    }
  std::vector<Sample> out;
  Rng rng(d.seed);
  const int nc[3] = {d.cells[0], dim > 1 ? d.cells[1] : 1, dim > 2 ? d.cells[2] : 1};
  const double inv_q = 1.0 / q;
  int c[3];
  for (c[0] = 0; c[0] < nc[0]; c[0]++)
    for (c[1] = 0; c[1] < nc[1]; c[1]++)
      for (c[2] = 0; c[2] < nc[2]; c[2]++)
        for (int sub = 0; sub < ppc; sub++) {
          Sample smp;
          smp.x[0] = smp.x[1] = smp.x[2] = 0;
          bool in_box = true;
          for (int a = 0; a < dim; a++) {
            const double j = d.jitter ? rng.u() : 0.5;
            smp.x[a] = d.lo[a] + (c[a] + (subs[(size_t)sub * 3 + a] + j) * inv_q) * h[a];
            in_box = in_box && smp.x[a] >= blo[a] && smp.x[a] <= bhi[a];
          }
          if (!in_box) continue;
          smp.shape = -1;
          for (int s = 0; s < d.nshapes; s++)
            if (inside(d, s, smp.x)) {
              smp.shape = s;
              break;
            }
          if (smp.shape >= 0) out.push_back(smp);
        }
  return out;
}

template<typename T>
void fill(const mpmx_seed_desc& d, const std::vector<Sample>& smp, T* X, T* V, T* F, T* B, T* mass, T* volume,
          T* mu, T* lambda, uint8_t* valid) {
  const int dim = d.dim;
  double cellvol = 1;
  for (int a = 0; a < dim; a++) cellvol *= (d.hi[a] - d.lo[a]) / d.cells[a];
  int ppc = 1;
  for (int a = 0; a < dim; a++) ppc *= d.ppc_axis;
  const double vol = cellvol / ppc;
  for (size_t p = 0; p < smp.size(); p++) {
    const int shape = smp[p].shape;
    const double E = d.youngs[shape], nu = d.poisson[shape];
    for (int a = 0; a < dim; a++) {
      X[p * dim + a] = (T)smp[p].x[a];
      V[p * dim + a] = (T)d.velocity[3 * shape + a];
    }
    for (int k = 0; k < dim * dim; k++) {
      F[p * dim * dim + k] = (k % (dim + 1) == 0) ? (T)1 : (T)0;
      B[p * dim * dim + k] = 0;
    }
    volume[p] = (T)vol;
    mass[p] = (T)(d.density[shape] * vol);
    mu[p] = (T)(E / (2 * (1 + nu)));  // STANDARD_TESTS_BASE.cpp:570-571
    lambda[p] = (T)(E * nu / ((1 + nu) * (1 - 2 * nu)));
    valid[p] = 1;
  }
}

}  // namespace

extern "C" int64_t mpmx_seed_particles(const mpmx_seed_desc* d, int precision, int64_t capacity, void* X, void* V,
                                       void* F, void* B, void* mass, void* volume, void* mu, void* lambda,
                                       uint8_t* valid) {
  if (!d || (d->dim != 2 && d->dim != 3) || d->ppc_axis < 1 || d->nshapes < 0) return -1;
  for (int a = 0; a < d->dim; a++)
    if (d->cells[a] < 1 || !(d->hi[a] > d->lo[a])) return -1;
  if (d->nshapes > 0 && (!d->shape_type || !d->shape_a || !d->shape_b || !d->density || !d->youngs ||
                         !d->poisson || !d->velocity))
    return -1;
  const std::vector<Sample> smp = seed_samples(*d);
  const int64_t n = (int64_t)smp.size();
  if (!X || capacity < n) return n;
  if (!V || !F || !B || !mass || !volume || !mu || !lambda || !valid) return -1;
  if (precision == 8)
    fill<double>(*d, smp, (double*)X, (double*)V, (double*)F, (double*)B, (double*)mass, (double*)volume,
                 (double*)mu, (double*)lambda, valid);
  else if (precision == 4)
    fill<float>(*d, smp, (float*)X, (float*)V, (float*)F, (float*)B, (float*)mass, (float*)volume, (float*)mu,
                (float*)lambda, valid);
  else
    return -1;
  return n;
}

extern "C" int mpmx_random_boxes(uint32_t seed, int nboxes, double smin, double smax, double rlo, double rhi,
                                 double Emin, double Emax, double numin, double numax, int max_tries, int cubes,
                                 double* box_lo, double* box_hi, double* E, double* nu) {
  if (nboxes < 0 || !(smax >= smin) || !(rhi > rlo) || !box_lo || !box_hi || !E || !nu) return -1;
  Rng rng(seed);
  int placed = 0;
  for (int b = 0; b < nboxes; b++) {
    bool ok = false;
    for (int t = 0; t < max_tries && !ok; t++) {
      double s[3], lo[3];
      if (cubes) {
        s[0] = s[1] = s[2] = rng.uniform(smin, smax);
      } else {
        for (int a = 0; a < 3; a++) s[a] = rng.uniform(smin, smax);
      }
      for (int a = 0; a < 3; a++) lo[a] = rng.uniform(rlo, rhi - s[a]);
      ok = true;
      for (int o = 0; o < placed && ok; o++) {
        bool sep = false;
        for (int a = 0; a < 3; a++)
          if (lo[a] + s[a] <= box_lo[3 * o + a] || box_hi[3 * o + a] <= lo[a]) sep = true;
        if (!sep) ok = false;
      }
      if (ok) {
        for (int a = 0; a < 3; a++) {
          box_lo[3 * placed + a] = lo[a];
          box_hi[3 * placed + a] = lo[a] + s[a];
        }
        E[placed] = std::exp(rng.uniform(std::log(Emin), std::log(Emax)));
        nu[placed] = rng.uniform(numin, numax);
        placed++;
      }
    }
    if (!ok) break;
  }
  return placed;
}
