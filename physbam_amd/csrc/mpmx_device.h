// Device math for the explicit APIC MPM step on gfx950: quadratic B-spline
// weights, a Jacobi-based SVD of F held in registers, and the per-particle
// Kirchhoff stress tau = V0 * P(F) * F^T used by the force scatter.
//
// Conventions follow PhysBAM's MPM path:
//   * grid nodes are MAC cell centres x_i = lo + (i + 1/2) h with three ghost
//     layers, so node coordinate n in [0, N+6) is reference index n-3
//     (Grid_Tools/Grids/GRID.h:123-124, :258-259);
//   * the stencil base is floor((x - h - lo)/h), local coordinate
//     r = (x - x_base)/h in [0.5, 1.5) (PARTICLE_GRID_WEIGHTS_SPLINE.cpp:133-140);
//   * the SVD returns rotations U, V with |s0|>=|s1|>=|s2| and only s2 possibly
//     negative (Core/Matrices/MATRIX_3X3.cpp:198-205, :115-122).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpmx {

#define MPMX_HD __host__ __device__ __forceinline__

template<typename T> MPMX_HD T tsqrt(T x);
template<> MPMX_HD float tsqrt<float>(float x) { return sqrtf(x); }
template<> MPMX_HD double tsqrt<double>(double x) { return sqrt(x); }
template<typename T> MPMX_HD T tabs(T x) { return x < 0 ? -x : x; }
template<typename T> MPMX_HD T tlog(T x);
template<> MPMX_HD float tlog<float>(float x) { return logf(x); }
template<> MPMX_HD double tlog<double>(double x) { return log(x); }
template<typename T> MPMX_HD T tfloor(T x);
template<> MPMX_HD float tfloor<float>(float x) { return floorf(x); }
template<> MPMX_HD double tfloor<double>(double x) { return floor(x); }

// Tile geometry: a tile is 4x4x4 nodes in 3D and 8x8 nodes in 2D (64 nodes).
// The LDS region of a particle tile covers the stencils of every particle whose
// base node lies in the tile: 6^3 (3D) or 10^2 (2D) nodes.
template<int D> struct Tile;
template<> struct Tile<3> { static constexpr int E = 4, R = 6, NODES = 64, RN = 216, K = 27; };
template<> struct Tile<2> { static constexpr int E = 8, R = 10, NODES = 64, RN = 100, K = 9; };

constexpr uint32_t INVALID_KEY = 0xffffffffu;
constexpr int MAX_WALLS = 16;

template<typename T> struct alignas(4 * sizeof(T)) Node4 { T m, x, y, z; };

template<typename T> struct WallDev {
  int type;  // 0 stick, 1 slip, 2 separate
  int pad;
  T friction;
  T lo[3], hi[3];
};

// Geometry of one scene's grid.  Several scenes share one tile array: scene
// tiles are numbered from tile0 on.
template<typename T, int D> struct Geom {
  T lo[D], hi[D], dx[D], inv_dx[D];
  T vlo[D], vhi[D];  // particles outside [vlo, vhi] are invalidated (MPM_DRIVER.cpp:1227-1238)
  int G[D];          // nodes per axis, N + 6
  int GT[D];         // tiles per axis (storage), ceil(G/E) + 1
  int ntiles;        // prod GT
  uint32_t tile0;    // first tile of this scene in the shared tile array
};

// Stencil base (node coordinate, ghost offset included) and 1D weights.
template<typename T, int D> struct Wts {
  int base[D];
  T xe[D];              // x - x_base
  T w[3][D], dw[3][D];
};

// Stencil base along one axis in node coordinates without the ghost offset:
// floor((x - dx - lo) / dx) (PARTICLE_GRID_WEIGHTS_SPLINE.cpp:33-56 with
// degree 2).  Every kernel that bins, classifies or weights a particle calls
// this, so they all agree on its cell.
template<typename T, int D>
MPMX_HD int stencil_base(const Geom<T, D>& g, const T* X, int a) {
  return (int)tfloor<T>(((X[a] - g.dx[a]) - g.lo[a]) * g.inv_dx[a]);
}

template<typename T, int D>
MPMX_HD void weights(const Geom<T, D>& g, const T* X, Wts<T, D>& s) {
#pragma unroll
  for (int a = 0; a < D; a++) {
    const T inv = g.inv_dx[a];
    int b = stencil_base<T, D>(g, X, a);
    T cen = g.lo[a] + ((T)b + (T)0.5) * g.dx[a];
    T xe = X[a] - cen;
    T x = xe * inv;
    s.base[a] = b + 3;
    s.xe[a] = xe;
    T q0 = x - (T)1.5, q1 = x - (T)1, q2 = x - (T)0.5;
    s.w[0][a] = (T)0.5 * q0 * q0;
    s.w[1][a] = (T)0.75 - q1 * q1;
    s.w[2][a] = (T)0.5 * q2 * q2;
    s.dw[0][a] = q0 * inv;
    s.dw[1][a] = (T)(-2) * q1 * inv;
    s.dw[2][a] = q2 * inv;
  }
}

// Binning key of a particle: id of the tile holding its stencil base node (in
// the shared tile array), or INVALID_KEY if the stencil would leave the ghosted
// grid.  Sorting by this key groups particles by tile.
template<typename T, int D>
MPMX_HD uint32_t tile_key(const Geom<T, D>& g, const T* X) {
  uint32_t tile = 0;
#pragma unroll
  for (int a = 0; a < D; a++) {
    const int b = stencil_base<T, D>(g, X, a) + 3;
    if (!(b >= 0 && b + 2 < g.G[a])) return INVALID_KEY;
    tile = tile * (uint32_t)g.GT[a] + (uint32_t)(b / Tile<D>::E);
  }
  return g.tile0 + tile;
}

// Label (0..63) of the cell with offsets o inside its tile, and back.  The
// kernels give lane l of a wavefront the cell labelled l, and each lane then
// reads or writes 16-byte nodes of a 6x6x6 region at 36 x + 6 y + z plus an
// offset common to all lanes.  3D labels therefore run z fastest, then x, then
// y: 16 consecutive labels (one y, x = 0..3, z = 0..3) fall on nodes
// 36 x + z apart, distinct modulo 16 since 36 = 4 (mod 16), so the 16 lanes
// that one LDS pass serves hit distinct banks (row-major labels put lanes
// 6 apart and collide in pairs).  2D labels are row-major.
template<int D> MPMX_HD uint32_t cell_label(const int* o) {
  if constexpr (D == 3) return (uint32_t)((o[1] * 4 + o[0]) * 4 + o[2]);
  else return (uint32_t)(o[0] * Tile<D>::E + o[1]);
}
template<int D> MPMX_HD void cell_offsets(uint32_t l, int* o) {
  if constexpr (D == 3) {
    o[2] = (int)(l & 3u);
    o[0] = (int)((l >> 2) & 3u);
    o[1] = (int)(l >> 4);
  } else {
    o[1] = (int)(l % Tile<D>::E);
    o[0] = (int)(l / Tile<D>::E);
  }
}

// -------------------------------------------------------------------- SVD --
// One-sided Jacobi: diagonalise S = F^T F with cyclic exact rotations (fixed
// sweep count), then B = F V, sort columns by norm, and Givens-QR B = U R.
// diag(R) are the singular values; R's off-diagonals vanish to rounding.

template<typename T> struct SweepCount;
template<> struct SweepCount<float> { static constexpr int n = 4; };
template<> struct SweepCount<double> { static constexpr int n = 6; };

// Rotation (c, s) annihilating S_pq of a symmetric 2x2 block [app apq; apq aqq].
template<typename T>
MPMX_HD void jacobi_rot(T app, T aqq, T apq, T& c, T& s) {
  if (apq == (T)0) { c = 1; s = 0; return; }
  T tau = (aqq - app) / ((T)2 * apq);
  T t = (tau >= 0 ? (T)1 : (T)-1) / (tabs(tau) + tsqrt(1 + tau * tau));
  c = (T)1 / tsqrt(1 + t * t);
  s = t * c;
}

// Column-major 3x3 (m[i + 3j]) SVD: F = U diag(s) V^T.
template<typename T>
MPMX_HD void svd3(const T* F, T* U, T* s, T* V) {
  // S = F^T F (symmetric)
  T S[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      S[i][j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
  T Vm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
  for (int sweep = 0; sweep < SweepCount<T>::n; sweep++) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int p = k == 2 ? 1 : 0, q = k == 0 ? 1 : 2;
      T c, sn;
      jacobi_rot(S[p][p], S[q][q], S[p][q], c, sn);
      // S <- J^T S J with J = I except J_pp=c, J_pq=s, J_qp=-s, J_qq=c
#pragma unroll
      for (int r = 0; r < 3; r++) {
        T sp = S[r][p], sq = S[r][q];
        S[r][p] = c * sp - sn * sq;
        S[r][q] = sn * sp + c * sq;
      }
#pragma unroll
      for (int r = 0; r < 3; r++) {
        T sp = S[p][r], sq = S[q][r];
        S[p][r] = c * sp - sn * sq;
        S[q][r] = sn * sp + c * sq;
      }
#pragma unroll
      for (int r = 0; r < 3; r++) {
        T vp = Vm[r][p], vq = Vm[r][q];
        Vm[r][p] = c * vp - sn * vq;
        Vm[r][q] = sn * vp + c * vq;
      }
    }
  }
  // B = F V (Bm[row][col])
  T Bm[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      Bm[i][j] = F[i] * Vm[0][j] + F[i + 3] * Vm[1][j] + F[i + 6] * Vm[2][j];
  T n0 = Bm[0][0] * Bm[0][0] + Bm[1][0] * Bm[1][0] + Bm[2][0] * Bm[2][0];
  T n1 = Bm[0][1] * Bm[0][1] + Bm[1][1] * Bm[1][1] + Bm[2][1] * Bm[2][1];
  T n2 = Bm[0][2] * Bm[0][2] + Bm[1][2] * Bm[1][2] + Bm[2][2] * Bm[2][2];
  // Sort columns by decreasing norm; a swap negates one column to keep det V = +1.
  auto swapc = [&](int a, int b, T& na, T& nb) {
#pragma unroll
    for (int r = 0; r < 3; r++) {
      T t = Bm[r][a]; Bm[r][a] = -Bm[r][b]; Bm[r][b] = t;
      T u = Vm[r][a]; Vm[r][a] = -Vm[r][b]; Vm[r][b] = u;
    }
    T t = na; na = nb; nb = t;
  };
  if (n0 < n1) swapc(0, 1, n0, n1);
  if (n0 < n2) swapc(0, 2, n0, n2);
  if (n1 < n2) swapc(1, 2, n1, n2);
  // Givens QR of B; U accumulates the rotations (det U = +1).
  T Um[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  auto givens = [&](int r0, int r1, int col) {
    T a = Bm[r0][col], b = Bm[r1][col];
    T rr = tsqrt(a * a + b * b);
    T c = 1, sn = 0;
    if (rr > (T)0) { c = a / rr; sn = b / rr; }
    // rows r0, r1 of B <- G B with G = [c s; -s c]; U <- U G^T
#pragma unroll
    for (int j = 0; j < 3; j++) {
      T x0 = Bm[r0][j], x1 = Bm[r1][j];
      Bm[r0][j] = c * x0 + sn * x1;
      Bm[r1][j] = -sn * x0 + c * x1;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
      T u0 = Um[i][r0], u1 = Um[i][r1];
      Um[i][r0] = c * u0 + sn * u1;
      Um[i][r1] = -sn * u0 + c * u1;
    }
  };
  givens(0, 1, 0);
  givens(0, 2, 0);
  givens(1, 2, 1);
  s[0] = Bm[0][0]; s[1] = Bm[1][1]; s[2] = Bm[2][2];
  // r00, r11 >= 0 by construction; only s[2] carries the sign of det F.
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) { U[i + 3 * j] = Um[i][j]; V[i + 3 * j] = Vm[i][j]; }
}

// Column-major 2x2 SVD with the same construction.
template<typename T>
MPMX_HD void svd2(const T* F, T* U, T* s, T* V) {
  T s00 = F[0] * F[0] + F[1] * F[1];
  T s11 = F[2] * F[2] + F[3] * F[3];
  T s01 = F[0] * F[2] + F[1] * F[3];
  T c, sn;
  jacobi_rot(s00, s11, s01, c, sn);
  T Vm[2][2] = {{c, sn}, {-sn, c}};
  T Bm[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) Bm[i][j] = F[i] * Vm[0][j] + F[i + 2] * Vm[1][j];
  T n0 = Bm[0][0] * Bm[0][0] + Bm[1][0] * Bm[1][0];
  T n1 = Bm[0][1] * Bm[0][1] + Bm[1][1] * Bm[1][1];
  if (n0 < n1) {
#pragma unroll
    for (int r = 0; r < 2; r++) {
      T t = Bm[r][0]; Bm[r][0] = -Bm[r][1]; Bm[r][1] = t;
      T u = Vm[r][0]; Vm[r][0] = -Vm[r][1]; Vm[r][1] = u;
    }
  }
  T a = Bm[0][0], b = Bm[1][0];
  T rr = tsqrt(a * a + b * b);
  T cc = 1, ss = 0;
  if (rr > (T)0) { cc = a / rr; ss = b / rr; }
  s[0] = cc * Bm[0][0] + ss * Bm[1][0];
  s[1] = -ss * Bm[0][1] + cc * Bm[1][1];
  U[0] = cc; U[1] = ss; U[2] = -ss; U[3] = cc;
  V[0] = Vm[0][0]; V[1] = Vm[1][0]; V[2] = Vm[0][1]; V[3] = Vm[1][1];
}

template<typename T, int D> MPMX_HD void svd(const T* F, T* U, T* s, T* V);
template<> MPMX_HD void svd<float, 3>(const float* F, float* U, float* s, float* V) { svd3(F, U, s, V); }
template<> MPMX_HD void svd<double, 3>(const double* F, double* U, double* s, double* V) { svd3(F, U, s, V); }
template<> MPMX_HD void svd<float, 2>(const float* F, float* U, float* s, float* V) { svd2(F, U, s, V); }
template<> MPMX_HD void svd<double, 2>(const double* F, double* U, double* s, double* V) { svd2(F, U, s, V); }

// ------------------------------------------------------------- stress ------
// Diagonal first Piola stress P_hat(sigma) in singular-value space.
//   fixed corotated: 2 mu (s - 1) + lambda (J - 1) cof(s)
//     (Deformables/Constitutive_Models/COROTATED_FIXED.cpp:59-65)
//   neo-Hookean:     mu s - (mu - lambda ln J) / s           for J >= t,
//     otherwise the hyperbola-clamped extrapolation of NEO_HOOKEAN.cpp:44-81.
//   St. Venant-Kirchhoff with Hencky strain: (2 mu ln s + lambda tr ln s) / s,
//     zero when the smallest singular value is not positive
//     (ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.cpp:60-70).
template<typename T, int D>
MPMX_HD void p_hat(int model, T mu, T la, T thr, T root, const T* s, T* P) {
  T J = s[0];
#pragma unroll
  for (int i = 1; i < D; i++) J *= s[i];
  if (model == 2) {
    if (!(s[D - 1] > (T)0)) {
#pragma unroll
      for (int i = 0; i < D; i++) P[i] = 0;
      return;
    }
    T l[D], tr = 0;
#pragma unroll
    for (int i = 0; i < D; i++) {
      l[i] = tlog<T>(s[i]);
      tr += l[i];
    }
#pragma unroll
    for (int i = 0; i < D; i++) P[i] = ((T)2 * mu * l[i] + la * tr) * ((T)1 / s[i]);
    return;
  }
  if (model == 0) {
    T cof[D];
    if constexpr (D == 3) { cof[0] = s[1] * s[2]; cof[1] = s[0] * s[2]; cof[2] = s[0] * s[1]; }
    else { cof[0] = s[1]; cof[1] = s[0]; }
#pragma unroll
    for (int i = 0; i < D; i++) P[i] = (T)2 * mu * (s[i] - (T)1) + la * (J - (T)1) * cof[i];
    return;
  }
  if (J >= thr) {
    T c = mu - la * tlog<T>(J);
#pragma unroll
    for (int i = 0; i < D; i++) P[i] = mu * s[i] - c / s[i];
    return;
  }
  // Clamp to the hyperbola prod(s) = t along the largest singular values.
  T Fc[D];
  if constexpr (D == 3) {
    if (s[0] * s[1] * s[1] > thr) { Fc[0] = s[0]; Fc[1] = s[1]; Fc[2] = thr / (s[0] * s[1]); }
    else if (s[0] * s[0] * s[0] > thr) { T q = tsqrt(thr / s[0]); Fc[0] = s[0]; Fc[1] = q; Fc[2] = q; }
    else { Fc[0] = root; Fc[1] = root; Fc[2] = root; }
  } else {
    if (s[0] * s[0] > thr) { Fc[0] = s[0]; Fc[1] = thr / s[0]; }
    else { Fc[0] = root; Fc[1] = root; }
  }
  T c = mu - la * tlog<T>(thr);
  T ip = 0, Fi[D], dF[D];
#pragma unroll
  for (int i = 0; i < D; i++) { Fi[i] = (T)1 / Fc[i]; dF[i] = s[i] - Fc[i]; ip += Fi[i] * dF[i]; }
#pragma unroll
  for (int i = 0; i < D; i++) P[i] = mu * s[i] + c * (Fi[i] * Fi[i] * dF[i] - Fi[i]) + la * ip * Fi[i];
}

// Energy density Psi of the singular values s (the term MPM_FINITE_ELEMENTS::
// Potential_Energy sums, Forces/MPM_FINITE_ELEMENTS.cpp:72-80):
//   fixed corotated: mu sum (s_i - 1)^2 + lambda/2 (J - 1)^2       (COROTATED_FIXED.cpp:179-185)
//   neo-Hookean:     mu (sum s_i^2 / 2 - d/2 - ln J) + lambda/2 ln^2 J, 1e16 for J <= 0
//                    (NEO_HOOKEAN.cpp:191-199; no clamp: the energy reads s itself)
//   Hencky:          mu sum ln^2 s_i + lambda/2 (sum ln s_i)^2, FLT_MAX if s_d <= 0
//                    (ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.cpp:49-59)
template<typename T, int D>
MPMX_HD T energy_density(int model, T mu, T la, const T* s) {
  T J = s[0];
#pragma unroll
  for (int i = 1; i < D; i++) J *= s[i];
  if (model == 0) {
    T q = 0;
#pragma unroll
    for (int i = 0; i < D; i++) q += (s[i] - (T)1) * (s[i] - (T)1);
    return mu * q + (T)0.5 * la * (J - (T)1) * (J - (T)1);
  }
  if (model == 1) {
    if (!(J > (T)0)) return (T)1e16;
    T i1 = 0;
#pragma unroll
    for (int i = 0; i < D; i++) i1 += s[i] * s[i];
    const T lj = tlog<T>(J);
    return mu * ((T)0.5 * (i1 - (T)D) - lj) + (T)0.5 * la * lj * lj;
  }
  if (!(s[D - 1] > (T)0)) return (T)3.402823466e38;
  T tr = 0, fn = 0;
#pragma unroll
  for (int i = 0; i < D; i++) {
    const T l = tlog<T>(s[i]);
    tr += l;
    fn += l * l;
  }
  return mu * fn + (T)0.5 * la * tr * tr;
}

// Sound speed of a particle (ISOTROPIC_CONSTITUTIVE_MODEL::Sound_Speed,
// ISOTROPIC_CONSTITUTIVE_MODEL.cpp:151-162) from its singular values s:
// sqrt(max(H_ii s_i^2, B_a s_{(a+1)%m}^2, B_a s_{(a+2)%m}^2) / density) with
// the diagonal of H and the spin entries B of Isotropic_Stress_Derivative
// (m = 1 in 2D, 3 in 3D):
//   fixed corotated  COROTATED_FIXED.cpp:127-140, :144-167 (panic threshold 1e-6, :22)
//   neo-Hookean      NEO_HOOKEAN.cpp:143-156, :161-180 (F clamped to the hyperbola below thr)
//   Hencky           ST_VENANT_KIRCHHOFF_HENCKY_STRAIN.cpp:85-102, :106-137 (failure threshold 0)
template<typename T> MPMX_HD T sound_panic(T d) {
  const T pt = (T)1e-6;
  return tabs(d) < pt ? (d < 0 ? -pt : pt) : d;
}
// (log x - log y) / (x - y) as Core/Math_Tools/Robust_Functions.h:33-39 forms it.
template<typename T> MPMX_HD T diff_log_over_diff(T x, T y) {
  const T z = (y - x) / x;
  const T r = tabs(z) < (T)1e-18 ? (T)1 : (T)log1p(z) / z;
  return r / x;
}
template<typename T, int D>
MPMX_HD T sound_speed(int model, T mu, T la, T thr, T root, const T* F, T density) {
  constexpr int M = D == 3 ? 3 : 1;
  T H[D], B[M];
  if (model == 0) {
    const T mu2 = 2 * mu;
    if constexpr (D == 3) {
      const T cof[3] = {F[1] * F[2], F[2] * F[0], F[1] * F[0]};
#pragma unroll
      for (int a = 0; a < 3; a++) H[a] = mu2 + la * (cof[a] * cof[a]);
      B[2] = mu2 * (1 - 1 / sound_panic(F[0] + F[1]));
      B[1] = mu2 * (1 - 1 / sound_panic(F[0] + F[2]));
      B[0] = mu2 * (1 - 1 / sound_panic(F[1] + F[2]));
    } else {
      H[0] = mu2 + la * (F[1] * F[1]);
      H[1] = mu2 + la * (F[0] * F[0]);
      B[0] = mu2 * (1 - 1 / sound_panic(F[0] + F[1]));
    }
  } else if (model == 1) {
    T J = F[0];
#pragma unroll
    for (int a = 1; a < D; a++) J *= F[a];
    T Fc[D];
    if (J >= thr) {
#pragma unroll
      for (int a = 0; a < D; a++) Fc[a] = F[a];
    } else if constexpr (D == 3) {
      if (F[0] * F[1] * F[1] > thr) { Fc[0] = F[0]; Fc[1] = F[1]; Fc[2] = thr / (F[0] * F[1]); }
      else if (F[0] * F[0] * F[0] > thr) { const T q = tsqrt(thr / F[0]); Fc[0] = F[0]; Fc[1] = q; Fc[2] = q; }
      else { Fc[0] = root; Fc[1] = root; Fc[2] = root; }
    } else {
      if (F[0] * F[0] > thr) { Fc[0] = F[0]; Fc[1] = thr / F[0]; }
      else { Fc[0] = root; Fc[1] = root; }
    }
    T Fi[D], ji = 1;
#pragma unroll
    for (int a = 0; a < D; a++) {
      Fi[a] = 1 / Fc[a];
      ji = a == 0 ? Fi[0] : ji * Fi[a];
    }
    const T mlj = mu + la * tlog<T>(ji);
#pragma unroll
    for (int a = 0; a < D; a++) H[a] = mu + (la + mlj) * (Fi[a] * Fi[a]);
#pragma unroll
    for (int a = 0; a < M; a++) B[a] = mu;
  } else {
    if (!(F[D - 1] > 0)) return 0;  // every entry of the derivative is zero
    T sv[D], l[D], sum_log = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
      sv[a] = F[a] > 0 ? F[a] : (T)0;
      l[a] = tlog<T>(sv[a]);
      sum_log += l[a];
    }
    const T lp2m = la + 2 * mu, two_mu = 2 * mu;
    if constexpr (D == 3) {
      T row[3];
#pragma unroll
      for (int i = 0; i < 3; i++) row[i] = sv[i] * l[0] + sv[i] * l[1] + sv[i] * l[2];
#pragma unroll
      for (int a = 0; a < 3; a++) H[a] = (lp2m - two_mu * l[a] - la * sum_log) / (sv[a] * sv[a]);
      const int pi[3] = {1, 0, 0}, pj[3] = {2, 2, 1};  // B(2): (0,1), B(1): (0,2), B(0): (1,2)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int i = pi[k], j = pj[k];
        const T xij = sv[j] * sv[i];
        const T mij = (two_mu / sv[i]) * (diff_log_over_diff(sv[i], sv[j]) - l[j] / sv[j]) - la * sum_log / xij;
        const T pij = (two_mu * (sv[j] * l[i] + sv[i] * l[j]) + la * (row[i] + row[j])) / (xij * (sv[i] + sv[j]));
        B[k] = (T)0.5 * (mij + pij);
      }
    } else {
      const T s0 = sv[0], s1 = sv[1], x10 = s1 * s0;
      const T m01 = (two_mu / s0) * (diff_log_over_diff(s0, s1) - l[1] / s1) - la * sum_log / x10;
      const T p01 = ((la * s0 + lp2m * s1) * l[0] + (la * s1 + lp2m * s0) * l[1]) / (x10 * (s0 + s1));
      H[0] = (lp2m * (1 - l[0]) - la * l[1]) / (s0 * s0);
      H[1] = (lp2m * (1 - l[1]) - la * l[0]) / (s1 * s1);
      B[0] = (T)0.5 * (m01 + p01);
    }
  }
  T F2[D];
#pragma unroll
  for (int a = 0; a < D; a++) F2[a] = F[a] * F[a];
  T cm = H[0] * F2[0];
#pragma unroll
  for (int a = 1; a < D; a++) cm = H[a] * F2[a] > cm ? H[a] * F2[a] : cm;
#pragma unroll
  for (int a = 0; a < M; a++)
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const T c = B[a] * F2[(a + i + 1) % M];
      cm = c > cm ? c : cm;
    }
  return tsqrt(cm / density);
}

// tau = vol * P F^T = vol * U diag(P_hat_i * s_i) U^T, symmetric; returned as a
// full column-major D x D matrix.  The FP64 path is used for near-degenerate F
// (singular values far apart in magnitude or the neo-Hookean clamp branch).
template<typename T, int D>
MPMX_HD void kirchhoff_typed(int model, T mu, T la, T thr, T root, T vol, const T* F, T* tau, bool& degenerate) {
  T U[D * D], V[D * D], s[D], P[D];
  svd<T, D>(F, U, s, V);
  T J = s[0];
#pragma unroll
  for (int i = 1; i < D; i++) J *= s[i];
  degenerate = (model == 1 && J < thr) || tabs(s[D - 1]) < (T)1e-3 * tabs(s[0]);
  p_hat<T, D>(model, mu, la, thr, root, s, P);
  T k[D];
#pragma unroll
  for (int i = 0; i < D; i++) k[i] = vol * P[i] * s[i];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T acc = 0;
#pragma unroll
      for (int q = 0; q < D; q++) acc += U[i + D * q] * k[q] * U[j + D * q];
      tau[i + D * j] = acc;
    }
}

template<typename T, int D>
__device__ __forceinline__ void kirchhoff(int model, T mu, T la, T thr, T root, T vol, const T* F, T* tau) {
  bool degenerate;
  kirchhoff_typed<T, D>(model, mu, la, thr, root, vol, F, tau, degenerate);
  if (sizeof(T) == 4 && degenerate) {
    double Fd[D * D], taud[D * D];
    bool dd;
#pragma unroll
    for (int i = 0; i < D * D; i++) Fd[i] = (double)F[i];
    kirchhoff_typed<double, D>(model, (double)mu, (double)la, (double)thr, (double)root, (double)vol, Fd, taud, dd);
#pragma unroll
    for (int i = 0; i < D * D; i++) tau[i] = (T)taud[i];
  }
}

// ------------------------------------------------------ fast stress path ---
// tau = vol * P(F) F^T evaluated from the small-strain quantities
//   G = F - I,  E = F F^T - I = G + G^T + G G^T,
//   J - 1 = tr G + (sum of principal 2x2 minors of G) + det G,
// which keeps full relative precision when F is close to a rotation.
//   neo-Hookean (J >= t):  tau = vol (mu E + lambda ln J I)          (no SVD)
//   fixed corotated:       tau = vol (2 mu ((I + E) - sqrt(I + E)) + lambda (J-1) J I)
//                          (F F^T - sqrt(F F^T) = (F - R) F^T for the polar R);
//     |E|_F <= r:          (1+x) - sqrt(1+x) by its Taylor series to x^7
//                          (corot_series; truncation below 2e-8 relative);
//     otherwise:           E = U diag(e) U^T by Jacobi; s_i = sqrt(1 + e_i);
//                          tau = vol U diag(2 mu s_i (s_i - 1) + lambda (J-1) J) U^T
//                          (FP32 3D: corot_closed_form, no eigenvectors)
// Inverted or nearly degenerate F (J <= 0 or s_min < 1e-2 s_max) and the
// neo-Hookean clamp branch (J < t) are declined (return false) and go through
// the FP64 SVD of kirchhoff_typed<double>().
template<typename T> __device__ __forceinline__ T fast_rcp(T x) { return (T)1 / x; }
template<> __device__ __forceinline__ float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template<typename T> __device__ __forceinline__ T fast_rsqrt(T x) { return (T)1 / sqrt(x); }
template<> __device__ __forceinline__ float fast_rsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template<typename T> __device__ __forceinline__ T fast_sqrt(T x) { return sqrt(x); }
template<> __device__ __forceinline__ float fast_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template<typename T> __device__ __forceinline__ T tlog1p(T x);
template<> __device__ __forceinline__ float tlog1p<float>(float x) { return log1pf(x); }
template<> __device__ __forceinline__ double tlog1p<double>(double x) { return log1p(x); }

template<typename T> struct EigenSweeps;
template<> struct EigenSweeps<float> { static constexpr int n = 4; };
template<> struct EigenSweeps<double> { static constexpr int n = 6; };

// Off-diagonal entries below this are left alone: a rotation for them moves
// nothing a strain of T's precision can hold, and 1 / a_pq of them would
// overflow (with equal diagonal entries, as for repeated stretches, the angle
// became 0 * inf = NaN).
template<typename T> struct TinyOffDiag;
template<> struct TinyOffDiag<float> { static constexpr float v = 1e-18f; };
template<> struct TinyOffDiag<double> { static constexpr double v = 1e-150; };

// Cyclic Jacobi on a symmetric D x D matrix A (in place); U collects the
// rotations so that A_in = U diag(A_out) U^T.
template<typename T, int D>
__device__ __forceinline__ void sym_eigen(T (&A)[D][D], T (&U)[D][D]) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) U[i][j] = i == j ? (T)1 : (T)0;
  constexpr int NP = D == 3 ? 3 : 1;
  constexpr int sweeps = D == 3 ? EigenSweeps<T>::n : 1;
#pragma unroll
  for (int sw = 0; sw < sweeps; sw++) {
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const int p = (D == 3 && k == 2) ? 1 : 0, q = (D == 3 && k != 0) ? 2 : 1;
      const T apq = A[p][q];
      if (tabs(apq) > TinyOffDiag<T>::v) {
        const T theta = (A[q][q] - A[p][p]) * (T)0.5 * fast_rcp<T>(apq);
        T t = fast_rcp<T>(tabs(theta) + fast_sqrt<T>((T)1 + theta * theta));
        t = theta < 0 ? -t : t;
        const T c = fast_rsqrt<T>((T)1 + t * t), s = t * c;
        A[p][p] -= t * apq;
        A[q][q] += t * apq;
        A[p][q] = A[q][p] = 0;
#pragma unroll
        for (int r = 0; r < D; r++) {
          if (r == p || r == q) continue;
          const T arp = A[r][p], arq = A[r][q];
          A[r][p] = A[p][r] = c * arp - s * arq;
          A[r][q] = A[q][r] = s * arp + c * arq;
        }
#pragma unroll
        for (int r = 0; r < D; r++) {
          const T up = U[r][p], uq = U[r][q];
          U[r][p] = c * up - s * uq;
          U[r][q] = s * up + c * uq;
        }
      }
    }
  }
}

// Largest |E|_F for the series path: |c_8| r^8 / (1 - r) stays below the
// precision of T relative to |E| / 2.
template<typename T> struct SeriesRadius;
template<> struct SeriesRadius<float> { static constexpr float r2 = 1.0f / 64; };
template<> struct SeriesRadius<double> { static constexpr double r2 = 1.0 / 65536; };

// (I + E) - sqrt(I + E) for symmetric E: Horner form of
//   x/2 + x^2/8 - x^3/16 + 5x^4/128 - 7x^5/256 + 21x^6/1024 - 33x^7/2048,
// the Taylor series of (1+x) - sqrt(1+x) (binomial series of sqrt(1+x)).
// Every partial product is a polynomial in E, so it commutes with E and is
// symmetric: only the upper triangle is formed.
// HENCKY: ln(I + E) = E - E^2/2 + E^3/3 - ... + E^7/7 instead (a template
// flag, so that the coefficients are immediates).
// DEG < 7 truncates the series at x^DEG (Horner from c_DEG).
template<typename T, int D, bool HENCKY, int DEG = 7>
__device__ __forceinline__ void corot_series(const T (&E)[D][D], T (&f)[D][D]) {
  const T c[8] = {(T)0,
                  HENCKY ? (T)1 : (T)0.5,
                  HENCKY ? (T)-0.5 : (T)0.125,
                  HENCKY ? (T)(1.0 / 3) : (T)-0.0625,
                  HENCKY ? (T)-0.25 : (T)(5.0 / 128),
                  HENCKY ? (T)0.2 : (T)(-7.0 / 256),
                  HENCKY ? (T)(-1.0 / 6) : (T)(21.0 / 1024),
                  HENCKY ? (T)(1.0 / 7) : (T)(-33.0 / 2048)};
  T P[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) P[i][j] = c[DEG] * E[i][j] + (i == j ? c[DEG - 1] : (T)0);
#pragma unroll
  for (int k = DEG - 2; k >= 0; k--) {
    T Q[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = i; j < D; j++) {
        T s = k > 0 && i == j ? c[k] : (T)0;
#pragma unroll
        for (int q = 0; q < D; q++) s += E[i][q] * P[q][j];
        Q[i][j] = s;
        Q[j][i] = s;
      }
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) P[i][j] = Q[i][j];
  }
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) f[i][j] = P[i][j];
}

// Fixed-corotated tau in 3D past the series radius, without eigenvectors.
// With b = F F^T = I + E and V = sqrt(b), the Cayley-Hamilton theorem gives
// (Hoger and Carlson 1984, "Determination of the stretch and rotation in the
// polar decomposition of the deformation gradient")
//   V = (-b^2 + (i1^2 - i2) b + i1 i3 I) / (i1 i2 - i3),  i_k the invariants of V.
// Written in the deviations d_k = s_k - 1 of the stretches (s_k^2 = 1 + e_k,
// e_k the eigenvalues of E) through their elementary symmetric functions
// q1, q2, q3, the corotated term is free of cancellation near b = I:
//   b - V = (c0 I + c1 E + E^2) / den,   c0 = -q3 (4 + q1),
//   c1 = 4 + 4 q1 + 3 q2 + q1^2 + q1 q2 - q3,
//   den = 8 + 8 q1 + 2 q2 + 2 q1^2 + q1 q2 - q3.
// The e_k come from the trigonometric solution of the characteristic cubic
// of E (O. K. Smith 1961, "Eigenvalues of a symmetric 3x3 matrix"); only
// symmetric functions of the d_k enter, so the accuracy lost in splitting a
// nearly double eigenvalue cancels.  Returns false for nearly singular F
// (s_min < 1e-2 s_max), as the eigen branch does.
template<typename T>
__device__ __forceinline__ bool corot_closed_form(const T (&E)[3][3], T mu, T pj, T vol, T* tau) {
  const T mE = (E[0][0] + E[1][1] + E[2][2]) * (T)(1.0 / 3);
  const T b00 = E[0][0] - mE, b11 = E[1][1] - mE, b22 = E[2][2] - mE;
  const T b01 = E[0][1], b02 = E[0][2], b12 = E[1][2];
  const T p2 = (b00 * b00 + b11 * b11 + b22 * b22 + (T)2 * (b01 * b01 + b02 * b02 + b12 * b12)) * (T)(1.0 / 6);
  T e1 = mE, e2 = mE, e3 = mE;  // e1 >= e2 >= e3
  if (p2 > (T)1e-30) {
    const T p = sqrt(p2);
    const T det = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) + b02 * (b01 * b12 - b11 * b02);
    T r = det / ((T)2 * p * p2);
    r = r < (T)-1 ? (T)-1 : (r > (T)1 ? (T)1 : r);
    const T phi = acos(r) * (T)(1.0 / 3);
    const T c = cos(phi), sn = sin(phi);
    e1 = mE + (T)2 * p * c;
    e3 = mE - p * (c + (T)1.7320508075688772 * sn);
    e2 = (T)3 * mE - e1 - e3;
  }
  if ((T)1 + e3 < (T)1e-4 * ((T)1 + e1)) return false;
  const T d1 = e1 / ((T)1 + sqrt((T)1 + e1)), d2 = e2 / ((T)1 + sqrt((T)1 + e2)), d3 = e3 / ((T)1 + sqrt((T)1 + e3));
  const T q1 = d1 + d2 + d3, q2 = d1 * d2 + d2 * d3 + d3 * d1, q3 = d1 * d2 * d3;
  const T c0 = -q3 * ((T)4 + q1);
  const T c1 = (T)4 + (T)4 * q1 + (T)3 * q2 + q1 * q1 + q1 * q2 - q3;
  const T inv = (T)1 / ((T)8 + (T)8 * q1 + (T)2 * q2 + (T)2 * q1 * q1 + q1 * q2 - q3);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = i; j < 3; j++) {
      T e2ij = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) e2ij += E[i][k] * E[k][j];
      const T bv = ((i == j ? c0 : (T)0) + c1 * E[i][j] + e2ij) * inv;
      const T t = vol * ((T)2 * mu * bv + (i == j ? pj : (T)0));
      tau[i + 3 * j] = t;
      tau[j + 3 * i] = t;
    }
  return true;
}

// Returns false (tau untouched) when F needs the FP64 SVD path.  CLOSED:
// FP32 3D fixed corotated past the series radius by corot_closed_form (else
// the Jacobi eigen branch, which needs fewer registers at once).
template<typename T, int D, bool CLOSED = true>
__device__ __forceinline__ bool kirchhoff_fast(int model, T mu, T la, T thr, T vol, const T* F, T* tau) {
  T G[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) G[i][j] = F[i + D * j] - (i == j ? (T)1 : (T)0);
  T E[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = i; j < D; j++) {
      T e = G[i][j] + G[j][i];
#pragma unroll
      for (int k = 0; k < D; k++) e += G[i][k] * G[j][k];
      E[i][j] = e;
      E[j][i] = e;
    }
  T Jm1;
  if constexpr (D == 3) {
    const T tr = G[0][0] + G[1][1] + G[2][2];
    const T m2 = (G[0][0] * G[1][1] - G[0][1] * G[1][0]) + (G[0][0] * G[2][2] - G[0][2] * G[2][0]) +
                 (G[1][1] * G[2][2] - G[1][2] * G[2][1]);
    const T dg = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                 G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    Jm1 = tr + m2 + dg;
  } else {
    Jm1 = G[0][0] + G[1][1] + (G[0][0] * G[1][1] - G[0][1] * G[1][0]);
  }
  const T J = (T)1 + Jm1;
  bool slow = J <= (T)0;
  if (model == 1) {
    if (!slow && J >= thr) {
      const T c = la * tlog1p<T>(Jm1);
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) tau[i + D * j] = vol * (mu * E[i][j] + (i == j ? c : (T)0));
      return true;
    }
    return false;
  }
  // fixed corotated (model 0) and Hencky (model 2) share the two branches:
  //   corotated tau = vol (2 mu ((I+E) - sqrt(I+E)) + lambda (J-1) J I),
  //   Hencky    tau = vol (mu ln(I+E) + lambda ln J I)  (= U diag(2 mu ln s + lambda tr ln s) U^T)
  const bool hencky = model == 2;
  if (!slow) {
    T r2 = 0;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) r2 += E[i][j] * E[i][j];
    // the log series converges more slowly: a quarter of the radius^2
    if (r2 <= (hencky ? SeriesRadius<T>::r2 * (T)0.25 : SeriesRadius<T>::r2)) {
      T f[D][D];
      if (hencky) {
        corot_series<T, D, true>(E, f);
      } else if constexpr (sizeof(T) == 4) {
        // FP32 fixed corotated: the series' relative truncation error after
        // x^K is about 2 |c_(K+1)| |E|^K, below 2e-8 for K = 7 at
        // |E|_F <= 1/8, for K = 5 at |E|_F <= 1/20 and for K = 3 at
        // |E|_F <= 1/160; the degree is the smallest that covers every lane
        // of the wavefront (a uniform branch)
        if (__any(r2 > (T)(1.0 / 400)))
          corot_series<T, D, false, 7>(E, f);
        else if (__any(r2 > (T)(1.0 / 25600)))
          corot_series<T, D, false, 5>(E, f);
        else
          corot_series<T, D, false, 3>(E, f);
      } else {
        corot_series<T, D, false>(E, f);
      }
      const T pj = hencky ? la * tlog1p<T>(Jm1) : la * Jm1 * J, m2 = hencky ? mu : (T)2 * mu;
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) tau[i + D * j] = vol * (m2 * f[i][j] + (i == j ? pj : (T)0));
      return true;
    }
    if constexpr (CLOSED && D == 3 && sizeof(T) == 4) {
      // FP32 fixed corotated: the closed form in place of the Jacobi sweeps
      if (!hencky) return corot_closed_form<T>(E, mu, la * Jm1 * J, vol, tau);
    }
    T U[D][D];
    sym_eigen<T, D>(E, U);
    T emin = E[0][0], emax = E[0][0];
#pragma unroll
    for (int i = 1; i < D; i++) {
      emin = E[i][i] < emin ? E[i][i] : emin;
      emax = E[i][i] > emax ? E[i][i] : emax;
    }
    if ((T)1 + emin < (T)1e-4 * ((T)1 + emax)) {
      slow = true;
    } else {
      const T pj = hencky ? la * tlog1p<T>(Jm1) : la * Jm1 * J;
      T k[D];
#pragma unroll
      for (int i = 0; i < D; i++) {
        const T e = E[i][i];
        if (hencky) {
          k[i] = vol * (mu * tlog1p<T>(e) + pj);
        } else {
          const T s = tsqrt<T>((T)1 + e);
          const T sm1 = e / ((T)1 + s);
          k[i] = vol * ((T)2 * mu * s * sm1 + pj);
        }
      }
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = i; j < D; j++) {
          T acc = 0;
#pragma unroll
          for (int q = 0; q < D; q++) acc += U[i][q] * k[q] * U[j][q];
          tau[i + D * j] = acc;
          tau[j + D * i] = acc;
        }
      return true;
    }
  }
  return false;
}

// FP64 SVD path for F that the fast path declines.  Matrices travel by value
// (in registers), so the out-of-line call needs no scratch memory.
template<typename T, int D> struct MatD {
  T v[D * D];
};

template<typename T, int D>
__device__ __noinline__ MatD<T, D> kirchhoff_slow_v(int model, T mu, T la, T thr, T root, T vol, MatD<T, D> F) {
  double Fd[D * D], taud[D * D];
  bool dd;
#pragma unroll
  for (int i = 0; i < D * D; i++) Fd[i] = (double)F.v[i];
  kirchhoff_typed<double, D>(model, (double)mu, (double)la, (double)thr, (double)root, (double)vol, Fd, taud, dd);
  MatD<T, D> tau;
#pragma unroll
  for (int i = 0; i < D * D; i++) tau.v[i] = (T)taud[i];
  return tau;
}

template<typename T, int D>
__device__ __forceinline__ void kirchhoff_slow(int model, T mu, T la, T thr, T root, T vol, const T* F, T* tau) {
  MatD<T, D> f;
#pragma unroll
  for (int i = 0; i < D * D; i++) f.v[i] = F[i];
  const MatD<T, D> t = kirchhoff_slow_v<T, D>(model, mu, la, thr, root, vol, f);
#pragma unroll
  for (int i = 0; i < D * D; i++) tau[i] = t.v[i];
}

template<typename T, int D>
__device__ __forceinline__ void kirchhoff_any(int model, T mu, T la, T thr, T root, T vol, const T* F, T* tau) {
  if (!kirchhoff_fast<T, D>(model, mu, la, thr, vol, F, tau)) kirchhoff_slow<T, D>(model, mu, la, thr, root, vol, F, tau);
}

// ---------------------------------------------------------------- walls ----
// Axis-aligned box signed distance and normal (Core/Math_Tools/RANGE.cpp:38-63).
template<typename T, int D>
MPMX_HD T box_phi(const WallDev<T>& w, const T* X) {
  T phi[D];
  bool inside = true;
#pragma unroll
  for (int i = 0; i < D; i++) {
    T c = (w.lo[i] + w.hi[i]) * (T)0.5, hl = (T)0.5 * (w.hi[i] - w.lo[i]);
    phi[i] = tabs(X[i] - c) - hl;
    if (!(phi[i] <= 0)) inside = false;
  }
  if (!inside) {
    T s = 0;
#pragma unroll
    for (int i = 0; i < D; i++) { T p = phi[i] > 0 ? phi[i] : (T)0; s += p * p; }
    return tsqrt(s);
  }
  T m = phi[0];
#pragma unroll
  for (int i = 1; i < D; i++) m = phi[i] > m ? phi[i] : m;
  return m;
}

template<typename T, int D>
MPMX_HD void box_normal(const WallDev<T>& w, const T* X, T* n) {
  bool inside = true;
  T pmax[D], pmin[D];
#pragma unroll
  for (int i = 0; i < D; i++) {
    if (!(X[i] >= w.lo[i] && X[i] <= w.hi[i])) inside = false;
    pmax[i] = X[i] - w.hi[i];
    pmin[i] = w.lo[i] - X[i];
    n[i] = 0;
  }
  if (inside) {
    // Face of least penetration; ties go to the lower axis.
    int axis = 0;
    T best = pmin[0] > pmax[0] ? pmin[0] : pmax[0];
#pragma unroll
    for (int i = 1; i < D; i++) {
      T m = pmin[i] > pmax[i] ? pmin[i] : pmax[i];
      if (m > best) { best = m; axis = i; }
    }
    n[axis] = pmax[axis] > pmin[axis] ? (T)1 : (T)-1;
    return;
  }
  T s = 0;
#pragma unroll
  for (int i = 0; i < D; i++) {
    T phi = pmin[i] > pmax[i] ? pmin[i] : pmax[i];
    n[i] = phi > 0 ? (pmax[i] > pmin[i] ? phi : -phi) : (T)0;
    s += n[i] * n[i];
  }
  T m = tsqrt(s);
  if (m > 0) {
    T inv = (T)1 / m;
#pragma unroll
    for (int i = 0; i < D; i++) n[i] *= inv;
  } else {
    n[0] = 1;
  }
}

}  // namespace mpmx
