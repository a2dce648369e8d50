// Adjoint (vector-Jacobian product) of one explicit APIC MPM step, the
// reverse of the step in mpmx_step.hip:
//
//   P2G   m_i = sum_p w m,  P_i = sum_p w m (v + C (x_i - x_p)) - dt tau grad w,
//         C = B D^-1, tau = V0 P(F) F^T
//   grid  u_i = P_i / m_i + dt g;  v_i = walls(u_i)  (MPM_OBJECTIVE.cpp:150-182)
//   G2P   v_p' = sum w v_i,  G = sum v_i grad w^T,  B' = sum w v_i (x_i - x_p)^T,
//         F' = (I + dt G) F,  x' = x + dt v_p'
//
// Given the adjoints (x', v', F', B')-bar of the new state, the kernels produce
// the adjoints of (x, v, F, B) and accumulate those of the per-particle Lame
// parameters in FP64:
//   g2p_adjoint  vt = v'bar + dt x'bar,  Gbar = dt F'bar F^T,
//                Fbar = (I + dt G)^T F'bar,
//                vbar_i += w vt + Gbar grad w + w B'bar (x_i - x_p)   (node scatter)
//                xbar   = x'bar + sum_i [(v_i.vt + v_i^T B'bar dz) grad w
//                                        + H (Gbar^T v_i) - w B'bar^T v_i]
//   grid_adjoint ubar = walls^T(vbar)  (stick: 0; slip/separate: (I - n n^T)),
//                Pbar = ubar / m,  mbar = -ubar.P / m^2
//   p2g_adjoint  vbar = m sum w Pbar,  Bbar = m sum w Pbar dz^T D^-1,
//                taubar = -dt sum Pbar grad w^T,
//                xbar += sum_i [(mbar m + m Pbar.(v + C dz)) grad w
//                               - m w C^T Pbar - dt H (tau Pbar)],
//                (Fbar, mubar, lambdabar) += d tau / d(F, mu, lambda) : taubar
// with dz = x_i - x_p and H the Hessian of w_ip in x_p (quadratic B-spline:
// second derivative (1, -2, 1) / h^2 per axis).  The derivative of the
// collision projection treats the face normal as constant, which is exact for
// face contacts (the normal of an inside point is a face normal).
#pragma once

#include "mpmx_device.h"

namespace mpmx {

// Node weight, gradient and Hessian of stencil entry k (row-major offsets).
template<typename T, int D>
__device__ __forceinline__ void stencil_entry(const Geom<T, D>& g, const Wts<T, D>& w, int k, int* o, T& wt, T* gr,
                                              T (*H)[D]) {
  int r = k;
#pragma unroll
  for (int c = D - 1; c >= 0; c--) {
    o[c] = r % 3;
    r /= 3;
  }
  T wa[D], da[D], ha[D];
#pragma unroll
  for (int a = 0; a < D; a++) {
    wa[a] = w.w[o[a]][a];
    da[a] = w.dw[o[a]][a];
    ha[a] = (o[a] == 1 ? (T)-2 : (T)1) * g.inv_dx[a] * g.inv_dx[a];
  }
  wt = 1;
#pragma unroll
  for (int a = 0; a < D; a++) wt *= wa[a];
#pragma unroll
  for (int a = 0; a < D; a++) {
    T ga = da[a];
#pragma unroll
    for (int b = 0; b < D; b++)
      if (b != a) ga *= wa[b];
    gr[a] = ga;
#pragma unroll
    for (int b = 0; b < D; b++) {
      T h = a == b ? ha[a] : da[a] * da[b];
#pragma unroll
      for (int c = 0; c < D; c++)
        if (c != a && c != b) h *= wa[c];
      H[a][b] = h;
    }
  }
}

// Adjoint of tau = vol P(F) F^T (column-major D x D) for the symmetric part S
// of taubar, in FP64:
//   neo-Hookean  tau = vol (mu (F F^T - I) + lambda ln J I)        (J >= threshold)
//     Fbar = vol (2 mu S F + lambda tr S F^-T)
//   fixed corotated  tau = vol (2 mu (F F^T - R F^T) + lambda (J-1) J I), R = U V^T
//     Fbar = vol (4 mu S F - 2 mu S R + lambda (2J-1) J tr S F^-T) + U A V^T,
//     A_ij = (M_ij - M_ji) / (s_i + s_j), M = U^T Rbar V, Rbar = -2 mu vol S F
//     (dR = U Omega V^T with Omega_ij = (U^T dF V)_ij - (U^T dF V)_ji over s_i + s_j,
//      for the SVD convention of MATRIX_3X3.cpp:198-236; |s_i + s_j| is floored
//      at 1e-6 as COROTATED_FIXED.cpp:150-152 does).
// Returns false for the neo-Hookean clamp branch (J < threshold), which has no
// adjoint here.
template<int D>
__device__ bool kirchhoff_adjoint(int model, double mu, double la, double thr, double vol, const double* F,
                                  const double* taubar, double* Fbar, double& mub, double& lab) {
  double S[D][D], Fm[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      S[i][j] = 0.5 * (taubar[i + D * j] + taubar[j + D * i]);
      Fm[i][j] = F[i + D * j];
    }
  double trS = 0;
#pragma unroll
  for (int i = 0; i < D; i++) trS += S[i][i];
  // J and F^-T from the cofactors
  double cof[D][D], J;
  if constexpr (D == 3) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        cof[i][j] = Fm[i1][j1] * Fm[i2][j2] - Fm[i1][j2] * Fm[i2][j1];
      }
    J = Fm[0][0] * cof[0][0] + Fm[0][1] * cof[0][1] + Fm[0][2] * cof[0][2];
  } else {
    cof[0][0] = Fm[1][1];
    cof[0][1] = -Fm[1][0];
    cof[1][0] = -Fm[0][1];
    cof[1][1] = Fm[0][0];
    J = Fm[0][0] * Fm[1][1] - Fm[0][1] * Fm[1][0];
  }
  // F^-T = cof / J
  double FFt[D][D], SF[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      double a = 0, b = 0;
#pragma unroll
      for (int k = 0; k < D; k++) {
        a += Fm[i][k] * Fm[j][k];
        b += S[i][k] * Fm[k][j];
      }
      FFt[i][j] = a;
      SF[i][j] = b;
    }
  if (model == 1) {
    if (!(J >= thr)) return false;
    double trSE = 0;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) trSE += S[i][j] * (FFt[i][j] - (i == j ? 1.0 : 0.0));
    mub = vol * trSE;
    lab = vol * log(J) * trS;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) Fbar[i + D * j] = vol * (2 * mu * SF[i][j] + la * trS * cof[i][j] / J);
    return true;
  }
  double U[D * D], s[D], V[D * D];
  svd<double, D>(F, U, s, V);
  double R[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      double r = 0;
#pragma unroll
      for (int k = 0; k < D; k++) r += U[i + D * k] * V[j + D * k];
      R[i][j] = r;
    }
  double trSB = 0;  // tr(S (F F^T - R F^T))
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      double rf = 0;
#pragma unroll
      for (int k = 0; k < D; k++) rf += R[i][k] * Fm[j][k];
      trSB += S[j][i] * (FFt[i][j] - rf);
    }
  mub = vol * 2 * trSB;
  lab = vol * (J - 1) * J * trS;
  // M = U^T Rbar V with Rbar = -2 mu vol S F
  double Rb[D][D], M[D][D], T1[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) Rb[i][j] = -2 * mu * vol * SF[i][j];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      double a = 0;
#pragma unroll
      for (int k = 0; k < D; k++) a += Rb[i][k] * V[k + D * j];
      T1[i][j] = a;  // Rbar V
    }
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      double a = 0;
#pragma unroll
      for (int k = 0; k < D; k++) a += U[k + D * i] * T1[k][j];
      M[i][j] = a;
    }
  double A[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      if (i == j) {
        A[i][j] = 0;
        continue;
      }
      double den = s[i] + s[j];
      if (fabs(den) < 1e-6) den = den < 0 ? -1e-6 : 1e-6;
      A[i][j] = (M[i][j] - M[j][i]) / den;
    }
  // Fbar = vol (4 mu S F - 2 mu S R + lambda (2J-1) J tr S F^-T) + U A V^T
  const double cj = la * (2 * J - 1) * trS;  // lambda (2J-1) J tr S times F^-T = cof / J
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      double sr = 0, uav = 0;
#pragma unroll
      for (int k = 0; k < D; k++) {
        sr += S[i][k] * R[k][j];
        double av = 0;
#pragma unroll
        for (int l = 0; l < D; l++) av += A[k][l] * V[j + D * l];
        uav += U[i + D * k] * av;
      }
      Fbar[i + D * j] = vol * (4 * mu * SF[i][j] - 2 * mu * sr + cj * cof[i][j]) + uav;
    }
  return true;
}

// The same adjoint in the simulation precision T (FP32 contexts), written on
// the forms the forward evaluates (kirchhoff_fast): with C = F F^T =
// I + E (E = G + G^T + G G^T, G = F - I, no cancellation) and S the symmetric
// part of taubar,
//   neo-Hookean      tau = vol (mu E + lambda ln J I)
//     Fbar = vol (2 mu S F + lambda tr S cof F / J),  mubar = vol S:E,
//     lambdabar = vol ln J tr S
//   fixed corotated  tau = vol (2 mu (C - sqrt C) + lambda (J - 1) J I)
//     Cbar = 2 mu vol (S - L(S)),  L(S) = Q [(Q^T S Q)_ij / (s_i + s_j)] Q^T
//     (the adjoint of d sqrt C, from sqrt C X + X sqrt C = dC in the eigenbasis
//     Q of E, s_i = sqrt(1 + e_i)),
//     Fbar = 2 Cbar F + vol lambda (2J - 1) tr S cof F,
//     mubar = 2 vol sum_i (Q^T S Q)_ii s_i (s_i - 1),  s_i - 1 = e_i / (1 + s_i),
//     lambdabar = vol (J - 1) J tr S.
// Returns 0 on success, 1 when F is inverted or nearly singular (J <= 0 or
// 1 + e_min < 1e-4 (1 + e_max): the FP64 SVD form kirchhoff_adjoint takes
// those), 2 for the neo-Hookean clamp branch (no adjoint).  mub and lab are
// returned in FP64 for the per-particle FP64 accumulation.
template<typename T, int D>
__device__ __forceinline__ int kirchhoff_adjoint_fast(int model, T mu, T la, T thr, T vol, const T* F, const T* taubar,
                                                      T* Fbar, double& mub, double& lab) {
  T S[D][D], Fm[D][D], G[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      S[i][j] = (T)0.5 * (taubar[i + D * j] + taubar[j + D * i]);
      Fm[i][j] = F[i + D * j];
      G[i][j] = Fm[i][j] - (i == j ? (T)1 : (T)0);
    }
  T trS = 0;
#pragma unroll
  for (int i = 0; i < D; i++) trS += S[i][i];
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
  T cof[D][D], Jm1;
  if constexpr (D == 3) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        cof[i][j] = Fm[i1][j1] * Fm[i2][j2] - Fm[i1][j2] * Fm[i2][j1];
      }
    const T tr = G[0][0] + G[1][1] + G[2][2];
    const T m2 = (G[0][0] * G[1][1] - G[0][1] * G[1][0]) + (G[0][0] * G[2][2] - G[0][2] * G[2][0]) +
                 (G[1][1] * G[2][2] - G[1][2] * G[2][1]);
    const T dg = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                 G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
    Jm1 = tr + m2 + dg;
  } else {
    cof[0][0] = Fm[1][1];
    cof[0][1] = -Fm[1][0];
    cof[1][0] = -Fm[0][1];
    cof[1][1] = Fm[0][0];
    Jm1 = G[0][0] + G[1][1] + (G[0][0] * G[1][1] - G[0][1] * G[1][0]);
  }
  const T J = (T)1 + Jm1;
  if (!(J > (T)0)) return 1;
  T SF[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T b = 0;
#pragma unroll
      for (int k = 0; k < D; k++) b += S[i][k] * Fm[k][j];
      SF[i][j] = b;
    }
  if (model == 1) {
    if (!(J >= thr)) return 2;
    T trSE = 0;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) trSE += S[i][j] * E[i][j];
    mub = (double)(vol * trSE);
    lab = (double)(vol * tlog1p<T>(Jm1) * trS);
    const T cl = la * trS / J;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) Fbar[i + D * j] = vol * ((T)2 * mu * SF[i][j] + cl * cof[i][j]);
    return 0;
  }
  // fixed corotated: eigenbasis of E
  T Ed[D][D], Q[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) Ed[i][j] = E[i][j];
  sym_eigen<T, D>(Ed, Q);
  T e[D], sg[D], emin = Ed[0][0], emax = Ed[0][0];
#pragma unroll
  for (int i = 0; i < D; i++) {
    e[i] = Ed[i][i];
    emin = e[i] < emin ? e[i] : emin;
    emax = e[i] > emax ? e[i] : emax;
  }
  if ((T)1 + emin < (T)1e-4 * ((T)1 + emax)) return 1;
#pragma unroll
  for (int i = 0; i < D; i++) sg[i] = tsqrt<T>((T)1 + e[i]);
  // W = Q^T S Q
  T SQ[D][D], W[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < D; k++) a += S[i][k] * Q[k][j];
      SQ[i][j] = a;
    }
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < D; k++) a += Q[k][i] * SQ[k][j];
      W[i][j] = a;
    }
  T mus = 0;
#pragma unroll
  for (int i = 0; i < D; i++) mus += W[i][i] * sg[i] * (e[i] / ((T)1 + sg[i]));
  mub = (double)((T)2 * vol * mus);
  lab = (double)(vol * Jm1 * J * trS);
  // K = Q [W_ij / (s_i + s_j)] Q^T;  Cbar = 2 mu vol (S - K)
  T Wd[D][D], KQ[D][D], Cb[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) Wd[i][j] = W[i][j] / (sg[i] + sg[j]);
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < D; k++) a += Q[i][k] * Wd[k][j];
      KQ[i][j] = a;
    }
  const T c2 = (T)2 * mu * vol;
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < D; k++) a += KQ[i][k] * Q[j][k];
      Cb[i][j] = c2 * (S[i][j] - a);
    }
  const T cj = vol * la * ((T)2 * J - (T)1) * trS;
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      T a = 0;
#pragma unroll
      for (int k = 0; k < D; k++) a += Cb[i][k] * Fm[k][j];
      Fbar[i + D * j] = (T)2 * a + cj * cof[i][j];
    }
  return 0;
}

}  // namespace mpmx
