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

// Poisson-disk sampling as STANDARD_TESTS_BASE::Seed_Particles_Poisson does
// it (Projects/mpm/STANDARD_TESTS_BASE.cpp:223-234) with POISSON_DISK<TV>
// (Geometry/Seeding/POISSON_DISK.cpp:29-127, Bridson's dart throwing over a
// background grid) and RANDOM_NUMBERS<T> (Core/Random_Numbers/
// RANDOM_NUMBERS.h:34-43), all in the precision T of the simulation so that
// the same seed gives the same points.
template<typename T, int D> struct PoissonDisk {
  std::mt19937 mt;
  T min_distance, h;
  int max_attempts = 30;
  static constexpr int ghost = 6;
  explicit PoissonDisk(uint32_t seed) : mt(seed) {}
  T number() { return (T)mt() * ((T)1 / (T)4294967296.0); }  // in [0, 1)
  T uniform(T a, T b) { return a + (b - a) * number(); }
  int uniform_integer(int a, int b) { return (int)(mt() % (unsigned long)(b + 1 - a)) + a; }
  // POISSON_DISK::Set_Distance_By_Volume
  void set_distance_by_volume(T v) {
    min_distance = D == 2 ? (T)std::sqrt(v * ((T)2 / 3)) : (T)std::cbrt(v * ((T)13 / 18));
    h = min_distance / (T)std::sqrt((T)D);
  }
  struct P {
    T x[D];
  };
  // Sample(random, box, X) for an empty X
  void sample(const T* lo, const T* hi, std::vector<P>& X) {
    const T th = min_distance * 2;
    T bmin[D], dmax[D], one_over_dx[D];
    int cc[D], span[D];
    for (int a = 0; a < D; a++) {
      bmin[a] = lo[a] - th;
      const T bmax = hi[a] + th;
      cc[a] = (int)((bmax - bmin[a]) / h) + 1;
      dmax[a] = bmin[a] + (T)cc[a] * h;  // GRID(cc + 1 nodes, [bmin, bmin + cc h]), not MAC
      const T dx = (dmax[a] - bmin[a]) / (T)cc[a];
      one_over_dx[a] = (T)1 / dx;
      span[a] = cc[a] + 2 * ghost;
    }
    size_t total = 1;
    for (int a = 0; a < D; a++) total *= (size_t)span[a];
    std::vector<int> cell_of(total, -1);  // the grid's cells including 6 ghost layers
    auto cell = [&](const P& p, int* c) {
      for (int a = 0; a < D; a++) c[a] = (int)std::floor((p.x[a] - bmin[a]) * one_over_dx[a]);
    };
    auto flat = [&](const int* c) {
      size_t f = 0;
      for (int a = 0; a < D; a++) f = f * (size_t)span[a] + (size_t)(c[a] + ghost);
      return f;
    };
    auto inside = [&](const P& p) {
      for (int a = 0; a < D; a++)
        if (!(p.x[a] >= lo[a] && p.x[a] <= hi[a])) return false;
      return true;
    };
    const T md2 = min_distance * min_distance;
    auto far_enough = [&](const P& p) {  // Check_Distance over the 5^D cells around p's cell
      int c[D], o[D];
      cell(p, c);
      for (int a = 0; a < D; a++) o[a] = -2;
      while (true) {
        int q[D];
        for (int a = 0; a < D; a++) q[a] = c[a] + o[a];
        const int k = cell_of[flat(q)];
        if (k != -1) {
          T m2 = 0;
          for (int a = 0; a < D; a++) {
            const T dd = p.x[a] - X[(size_t)k].x[a];
            m2 = a == 0 ? dd * dd : m2 + dd * dd;
          }
          if (m2 < md2) return false;
        }
        int a = D - 1;
        while (a >= 0 && ++o[a] == 3) o[a--] = -2;
        if (a < 0) return true;
      }
    };
    P first;
    for (int a = 0; a < D; a++) first.x[a] = uniform(lo[a], hi[a]);
    int c0[D];
    cell(first, c0);
    cell_of[flat(c0)] = 0;
    X.push_back(first);
    std::vector<int> active(1, 0);
    while (!active.empty()) {
      const int ri = uniform_integer(0, (int)active.size() - 1);
      const P point = X[(size_t)active[(size_t)ri]];
      bool found = false;
      for (int i = 0; i < max_attempts; i++) {
        P np;
        while (true) {  // Generate_Random_Point_Around_Annulus
          T v[D], m2 = 0;
          for (int a = 0; a < D; a++) v[a] = uniform((T)-1, (T)1);
          for (int a = 0; a < D; a++) m2 = a == 0 ? v[a] * v[a] : m2 + v[a] * v[a];
          if (m2 >= 0.25 && m2 <= 1) {
            for (int a = 0; a < D; a++) np.x[a] = v[a] * min_distance * 2 + point.x[a];
            break;
          }
        }
        if (far_enough(np) && inside(np)) {
          found = true;
          const int index = (int)X.size();
          X.push_back(np);
          active.push_back(index);
          int cn[D];
          cell(np, cn);
          cell_of[flat(cn)] = index;
          break;
        }
      }
      if (!found) {  // Remove_Index_Lazy
        active[(size_t)ri] = active.back();
        active.pop_back();
      }
    }
  }
};

template<typename T, int D>
std::vector<Sample> poisson_samples(const mpmx_seed_desc& d, double ppc) {
  PoissonDisk<T, D> pd(d.seed);
  T vol = 1;
  for (int a = 0; a < D; a++) {
    const T dx = ((T)d.hi[a] - (T)d.lo[a]) / (T)d.cells[a];  // MAC grid dX
    vol = a == 0 ? dx : vol * dx;
  }
  pd.set_distance_by_volume(vol / (T)ppc);
  std::vector<Sample> out;
  for (int s = 0; s < d.nshapes; s++) {
    T lo[D], hi[D];
    const double* a = d.shape_a + 3 * s;
    const double* b = d.shape_b + 3 * s;
    for (int k = 0; k < D; k++) {
      // the implicit object's box: the range, or the sphere's bounding box
      lo[k] = d.shape_type[s] == MPMX_SHAPE_BOX ? (T)a[k] : (T)a[k] - (T)b[0];
      hi[k] = d.shape_type[s] == MPMX_SHAPE_BOX ? (T)b[k] : (T)a[k] + (T)b[0];
    }
    std::vector<typename PoissonDisk<T, D>::P> X;
    pd.sample(lo, hi, X);
    if (d.shape_type[s] == MPMX_SHAPE_SPHERE)  // drop points with Extended_Phi > 0, last swapped in
      for (int i = (int)X.size() - 1; i >= 0; i--) {
        T m2 = 0;
        for (int k = 0; k < D; k++) {
          const T dd = X[(size_t)i].x[k] - (T)a[k];
          m2 = k == 0 ? dd * dd : m2 + dd * dd;
        }
        if ((T)std::sqrt(m2) - (T)b[0] > 0) {
          X[(size_t)i] = X.back();
          X.pop_back();
        }
      }
    for (const auto& p : X) {
      Sample smp;
      smp.x[0] = smp.x[1] = smp.x[2] = 0;
      for (int k = 0; k < D; k++) smp.x[k] = (double)p.x[k];
      smp.shape = s;
      out.push_back(smp);
    }
  }
  return out;
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

namespace {
// Seed_Particles_Poisson's particle volume and mass in T: volume =
// dX.Product() / particles_per_cell, mass = density * volume (:230-231).
template<typename T> void poisson_mass(const mpmx_seed_desc& d, double ppc, const std::vector<Sample>& smp, T* mass,
                                       T* volume) {
  T vol = 1;
  for (int a = 0; a < d.dim; a++) {
    const T dx = ((T)d.hi[a] - (T)d.lo[a]) / (T)d.cells[a];
    vol = a == 0 ? dx : vol * dx;
  }
  vol = vol / (T)ppc;
  for (size_t p = 0; p < smp.size(); p++) {
    volume[p] = vol;
    mass[p] = (T)d.density[smp[p].shape] * vol;
  }
}
}  // namespace

extern "C" int64_t mpmx_seed_particles_poisson(const mpmx_seed_desc* d, int precision, double particles_per_cell,
                                               int64_t capacity, void* X, void* V, void* F, void* B, void* mass,
                                               void* volume, void* mu, void* lambda, uint8_t* valid) {
  if (!d || (d->dim != 2 && d->dim != 3) || d->nshapes < 0 || (precision != 4 && precision != 8)) return -1;
  for (int a = 0; a < d->dim; a++)
    if (d->cells[a] < 1 || !(d->hi[a] > d->lo[a])) return -1;
  if (d->nshapes > 0 && (!d->shape_type || !d->shape_a || !d->shape_b || !d->density || !d->youngs ||
                         !d->poisson || !d->velocity))
    return -1;
  for (int s = 0; s < d->nshapes; s++) {  // a non-empty shape, so the dart throwing terminates
    if (d->shape_type[s] == MPMX_SHAPE_SPHERE && !(d->shape_b[3 * s] > 0)) return -1;
    if (d->shape_type[s] == MPMX_SHAPE_BOX)
      for (int a = 0; a < d->dim; a++)
        if (!(d->shape_b[3 * s + a] >= d->shape_a[3 * s + a])) return -1;
  }
  double ppc = particles_per_cell;
  if (!(ppc > 0)) {
    if (d->ppc_axis < 1) return -1;
    ppc = 1;
    for (int a = 0; a < d->dim; a++) ppc *= d->ppc_axis;  // the reference default 2^d for ppc_axis 2
  }
  std::vector<Sample> smp;
  if (precision == 8)
    smp = d->dim == 2 ? poisson_samples<double, 2>(*d, ppc) : poisson_samples<double, 3>(*d, ppc);
  else
    smp = d->dim == 2 ? poisson_samples<float, 2>(*d, ppc) : poisson_samples<float, 3>(*d, ppc);
  const int64_t n = (int64_t)smp.size();
  if (!X || capacity < n) return n;
  if (!V || !F || !B || !mass || !volume || !mu || !lambda || !valid) return -1;
  if (precision == 8) {
    fill<double>(*d, smp, (double*)X, (double*)V, (double*)F, (double*)B, (double*)mass, (double*)volume,
                 (double*)mu, (double*)lambda, valid);
    poisson_mass<double>(*d, ppc, smp, (double*)mass, (double*)volume);
  } else {
    fill<float>(*d, smp, (float*)X, (float*)V, (float*)F, (float*)B, (float*)mass, (float*)volume, (float*)mu,
                (float*)lambda, valid);
    poisson_mass<float>(*d, ppc, smp, (float*)mass, (float*)volume);
  }
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
