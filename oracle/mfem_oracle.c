/*
 * mfem_oracle.c -- CPU restatement of warp-mfem's per-timestep soft-body solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Built twice by oracle/Makefile: REAL=double (the FP64 oracle) and REAL=float
 * (same algorithm at the reference's FP32 precision).
 *
 * What it restates (paths relative to the reference repository):
 *   step order              src/mfem/refinement/solver.py:429-579
 *   predictor / velocity    src/mfem/refinement/kernels.py:77-91, 351-363
 *   mass + attachments      src/mfem/refinement/kernels.py:7-30, solver.py:320-393,
 *                           src/mfem/boundary_condition.py:10-92
 *   kinetic gradient        src/mfem/refinement/kernels.py:172-185
 *   constraints c, G        src/mfem/refinement/kernels.py:131-169,
 *                           src/mfem/refinement/utils.py:46-177
 *   ARAP / Neo-Hookean      src/mfem/refinement/elastic_energy.py:106-407,
 *                           src/mfem/utils.py:56-97
 *   GMW81 + PSD fix         src/mfem/numerical/gmw81.py:18-86, psd_fix.py:24-105
 *   Schur blocks            src/mfem/refinement/kernels.py:189-249
 *   assembly                src/mfem/refinement/solver.py:641-799
 *   block Jacobi            src/mfem/refinement/kernels.py:32-53, solver.py:581-609
 *   PCG                     warp.optim.linear.cg (warp-lang 1.15.0, not vendored):
 *                           textbook preconditioned CG, stop at |r| <= tol |b|
 *   local solve             src/mfem/refinement/kernels.py:252-283
 *   line search             src/mfem/line_search.py:12-163, solver.py:803-928,
 *                           src/mfem/refinement/kernels.py:285-348
 *   vertex contact/friction src/mfem/refinement/contact.py:16-387, 778-1106,
 *                           src/mfem/ipc/distance.py:4-57
 *   triangle vs capsule     src/mfem/refinement/contact.py:417-775,
 *                           src/mfem/collision.py:276-322
 *   membrane shell          src/mfem/refinement/shell_energy.py:37-331,
 *                           solver.py:291-302, 638-639, 789-797, 864-867, 910-911
 *
 * Functions with no check in tests/ that pins them are listed as
 * "parity unpinned" in the header comment of tests/test_oracle.py; the pins
 * of friction, shape distances, the triangle barrier, the local solve and
 * the GMW81 correction are in tests/test_oracle_pins.py, those of the
 * membrane shell in tests/test_oracle_shell.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif
typedef REAL real;

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* small dense algebra (row-major)                                      */
/* ------------------------------------------------------------------ */
static void mm(const real* A, const real* B, real* C, int n, int k, int m) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) {
      real acc = 0;
      for (int l = 0; l < k; ++l) acc += A[i * k + l] * B[l * m + j];
      C[i * m + j] = acc;
    }
}
static real det3(const real* A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}
static void inv3(const real* A, real* B) {
  real d = det3(A);
  real id = (real)1 / d;
  B[0] = (A[4] * A[8] - A[5] * A[7]) * id;
  B[1] = (A[2] * A[7] - A[1] * A[8]) * id;
  B[2] = (A[1] * A[5] - A[2] * A[4]) * id;
  B[3] = (A[5] * A[6] - A[3] * A[8]) * id;
  B[4] = (A[0] * A[8] - A[2] * A[6]) * id;
  B[5] = (A[2] * A[3] - A[0] * A[5]) * id;
  B[6] = (A[3] * A[7] - A[4] * A[6]) * id;
  B[7] = (A[1] * A[6] - A[0] * A[7]) * id;
  B[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}
static real dot3(const real* a, const real* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void cross3(const real* a, const real* b, real* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static real norm3(const real* a) { return (real)sqrt((double)dot3(a, a)); }

/* ------------------------------------------------------------------ */
/* 3x3 SVD, rotation variant                                           */
/* Jacobi eigen-decomposition of F^T F, descending sort, QR of F V      */
/* (the construction of McAdams et al. 2011, "Computing the SVD of     */
/* 3x3 matrices with minimal branching", done here with plain loops). */
/* U and V come out as rotations and the sign of det F lands on the    */
/* smallest singular value, which is the convention rotational_svd     */
/* establishes (src/mfem/refinement/utils.py:46-67).                   */
/* ------------------------------------------------------------------ */
static void jacobi_eig3(real* A, real* V) {
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1 : 0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    real off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    real dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= (real)1e-30 * dia || off == 0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        real apq = A[p * 3 + q];
        if (apq == 0) continue;
        real app = A[p * 3 + p], aqq = A[q * 3 + q];
        real theta = (aqq - app) / (2 * apq);
        real t = (theta >= 0 ? (real)1 : (real)-1) / ((real)fabs((double)theta) + (real)sqrt((double)(theta * theta + 1)));
        real cs = (real)1 / (real)sqrt((double)(t * t + 1)), sn = t * cs;
        for (int k = 0; k < 3; ++k) { /* A <- J^T A J */
          real akp = A[k * 3 + p], akq = A[k * 3 + q];
          A[k * 3 + p] = cs * akp - sn * akq;
          A[k * 3 + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 3; ++k) {
          real apk = A[p * 3 + k], aqk = A[q * 3 + k];
          A[p * 3 + k] = cs * apk - sn * aqk;
          A[q * 3 + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          real vkp = V[k * 3 + p], vkq = V[k * 3 + q];
          V[k * 3 + p] = cs * vkp - sn * vkq;
          V[k * 3 + q] = sn * vkp + cs * vkq;
        }
      }
  }
}

EXPORT void orc_svd3(const real* F, real* U, real* sig, real* V) {
  real A[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      real acc = 0;
      for (int k = 0; k < 3; ++k) acc += F[k * 3 + i] * F[k * 3 + j];
      A[i * 3 + j] = acc;
    }
  jacobi_eig3(A, V);
  real ev[3] = {A[0], A[4], A[8]};
  /* sort descending, permuting V's columns */
  for (int a = 0; a < 2; ++a)
    for (int b = a + 1; b < 3; ++b)
      if (ev[b] > ev[a]) {
        real t = ev[a]; ev[a] = ev[b]; ev[b] = t;
        for (int k = 0; k < 3; ++k) { t = V[k * 3 + a]; V[k * 3 + a] = V[k * 3 + b]; V[k * 3 + b] = t; }
      }
  if (det3(V) < 0)
    for (int k = 0; k < 3; ++k) V[k * 3 + 2] = -V[k * 3 + 2];
  real B[9];
  mm(F, V, B, 3, 3, 3);
  real b0[3] = {B[0], B[3], B[6]}, b1[3] = {B[1], B[4], B[7]}, b2[3] = {B[2], B[5], B[8]};
  real u0[3], u1[3], u2[3];
  real n0 = norm3(b0);
  if (n0 > (real)1e-30) { u0[0] = b0[0] / n0; u0[1] = b0[1] / n0; u0[2] = b0[2] / n0; }
  else { u0[0] = 1; u0[1] = 0; u0[2] = 0; }
  real pr = dot3(u0, b1);
  real w[3] = {b1[0] - pr * u0[0], b1[1] - pr * u0[1], b1[2] - pr * u0[2]};
  real n1 = norm3(w);
  if (n1 > (real)1e-12 * (n0 > 0 ? n0 : 1)) { u1[0] = w[0] / n1; u1[1] = w[1] / n1; u1[2] = w[2] / n1; }
  else {
    int ax = 0;
    real m = (real)fabs((double)u0[0]);
    if ((real)fabs((double)u0[1]) < m) { ax = 1; m = (real)fabs((double)u0[1]); }
    if ((real)fabs((double)u0[2]) < m) ax = 2;
    real e[3] = {0, 0, 0};
    e[ax] = 1;
    cross3(u0, e, u1);
    real nn = norm3(u1);
    u1[0] /= nn; u1[1] /= nn; u1[2] /= nn;
  }
  cross3(u0, u1, u2);
  sig[0] = dot3(u0, b0);
  sig[1] = dot3(u1, b1);
  sig[2] = dot3(u2, b2);
  for (int k = 0; k < 3; ++k) { U[k * 3 + 0] = u0[k]; U[k * 3 + 1] = u1[k]; U[k * 3 + 2] = u2[k]; }
}

/* ------------------------------------------------------------------ */
/* per-tet kinematics: refinement/kernels.py:131-169, utils.py:70-177   */
/* ------------------------------------------------------------------ */
static const real SYMD[6] = {1, 1, 1, 2, 2, 2};

/* 9x9 dR/dF, column 3j+i is the column-major flatten of dR/dF_ij
   (refinement/utils.py:93-121). */
EXPORT void orc_rotation_gradient(const real* U, const real* sig, const real* V, real* dRdF) {
  static const real T[3][9] = {
      {0, -1, 0, 1, 0, 0, 0, 0, 0},  /* twist in the (0,1) plane */
      {0, 0, 0, 0, 0, 1, 0, -1, 0},  /* (1,2) */
      {0, 0, 1, 0, 0, 0, -1, 0, 0}}; /* (0,2) */
  real s01 = sig[0] + sig[1], s12 = sig[1] + sig[2], s02 = sig[0] + sig[2];
  if (s01 < (real)1e-8) s01 = (real)1e-8;
  if (s12 < (real)1e-8) s12 = (real)1e-8;
  if (s02 < (real)1e-8) s02 = (real)1e-8;
  real coef[3] = {2 / s01, 2 / s12, 2 / s02};
  real t[3][9];
  real Vt[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Vt[i * 3 + j] = V[j * 3 + i];
  real isq = (real)(1.0 / sqrt(2.0));
  for (int k = 0; k < 3; ++k) {
    real UT[9], M[9];
    mm(U, T[k], UT, 3, 3, 3);
    mm(UT, Vt, M, 3, 3, 3);
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) t[k][3 * c + r] = isq * M[r * 3 + c];
  }
  for (int a = 0; a < 9; ++a)
    for (int b = 0; b < 9; ++b)
      dRdF[a * 9 + b] = coef[0] * t[0][a] * t[0][b] + coef[1] * t[1][a] * t[1][b] + coef[2] * t[2][a] * t[2][b];
}

/* 6x9 dS/dF, column 3i+j is d sym6(S)/dF_ij (refinement/utils.py:75-90). */
EXPORT void orc_stretch_gradient(const real* F, const real* U, const real* sig, const real* V, real* dSdF) {
  real dRdF[81];
  orc_rotation_gradient(U, sig, V, dRdF);
  real R[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      real acc = 0;
      for (int k = 0; k < 3; ++k) acc += U[i * 3 + k] * V[j * 3 + k];
      R[i * 3 + j] = acc;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      real dR[9], dS[9];
      int col = 3 * j + i;
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) dR[r * 3 + c] = dRdF[(3 * c + r) * 9 + col];
      /* dS = F^T dR */
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          real acc = 0;
          for (int k = 0; k < 3; ++k) acc += F[k * 3 + r] * dR[k * 3 + c];
          dS[r * 3 + c] = acc;
        }
      for (int c = 0; c < 3; ++c) dS[j * 3 + c] += R[i * 3 + c];
      int out = 3 * i + j;
      dSdF[0 * 9 + out] = dS[0];
      dSdF[1 * 9 + out] = dS[4];
      dSdF[2 * 9 + out] = dS[8];
      dSdF[3 * 9 + out] = dS[1];
      dSdF[4 * 9 + out] = dS[2];
      dSdF[5 * 9 + out] = dS[5];
    }
}

/* 9x12 dF/dx, row 3i+j <-> F[i,j], column 3k+i <-> vertex k coordinate i
   (refinement/utils.py:123-177). */
static void dFdx_from_dminv(const real* D, real* out) {
  memset(out, 0, sizeof(real) * 108);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      int row = 3 * i + j;
      real sum = D[0 * 3 + j] + D[1 * 3 + j] + D[2 * 3 + j];
      out[row * 12 + 0 * 3 + i] = -sum;
      for (int k = 0; k < 3; ++k) out[row * 12 + (k + 1) * 3 + i] = D[k * 3 + j];
    }
}

static void deformation_gradient(const real* x, const real* D, real* F) {
  real Ds[9];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) Ds[i * 3 + k] = x[(k + 1) * 3 + i] - x[i];
  mm(Ds, D, F, 3, 3, 3);
}

/* c = SYM (sym6(S) - s), G = -SYM dS/dF dF/dx (6x12 row-major), S6 = sym6(S). */
EXPORT void orc_tet_constraints(const real* x, const real* D, const real* s, real* c, real* G, real* S6) {
  real F[9], U[9], sg[3], V[9];
  deformation_gradient(x, D, F);
  orc_svd3(F, U, sg, V);
  real S[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      real acc = 0;
      for (int k = 0; k < 3; ++k) acc += V[i * 3 + k] * sg[k] * V[j * 3 + k];
      S[i * 3 + j] = acc;
    }
  real sF[6] = {S[0], S[4], S[8], S[1], S[2], S[5]};
  if (S6) for (int k = 0; k < 6; ++k) S6[k] = sF[k];
  for (int k = 0; k < 6; ++k) c[k] = SYMD[k] * (sF[k] - s[k]);
  if (G) {
    real dSdF[54], dFdx[108];
    orc_stretch_gradient(F, U, sg, V, dSdF);
    dFdx_from_dminv(D, dFdx);
    mm(dSdF, dFdx, G, 6, 9, 12);
    for (int r = 0; r < 6; ++r)
      for (int k = 0; k < 12; ++k) G[r * 12 + k] *= -SYMD[r];
  }
}

/* ------------------------------------------------------------------ */
/* GMW81 modified Cholesky and PSD fix (numerical/gmw81.py, psd_fix.py) */
/* ------------------------------------------------------------------ */
EXPORT void orc_gmw81(const real* Ain, int n, real* L) {
  real A[36];
  memcpy(A, Ain, sizeof(real) * n * n);
  const real eps = (real)1.19e-7, delta = (real)sqrt(1.19e-7);
  real gamma = 0, xi = 0;
  for (int i = 0; i < n; ++i) {
    real v = (real)fabs((double)A[i * n + i]);
    if (v > gamma) gamma = v;
    for (int j = i + 1; j < n; ++j) {
      v = (real)fabs((double)A[i * n + j]);
      if (v > xi) xi = v;
    }
  }
  real b2 = xi / (real)sqrt((double)(n * n - 1));
  if (eps > b2) b2 = eps;
  if (gamma > b2) b2 = gamma;
  for (int j = 0; j < n; ++j) {
    for (int i = j + 1; i < n; ++i) {
      real cij = A[j * n + i];
      for (int k = 0; k < j; ++k) cij -= A[k * n + k] * A[k * n + k] * A[i * n + k] * A[j * n + k];
      A[i * n + j] = cij;
    }
    real theta = 0;
    for (int i = j + 1; i < n; ++i) {
      real v = (real)fabs((double)A[i * n + j]);
      if (v > theta) theta = v;
    }
    real cjj = A[j * n + j];
    real dj = (real)fabs((double)cjj);
    if (theta * theta / b2 > dj) dj = theta * theta / b2;
    if (delta > dj) dj = delta;
    A[j * n + j] = (real)sqrt((double)dj);
    for (int i = j + 1; i < n; ++i) {
      real cij = A[i * n + j];
      A[i * n + j] = cij / dj;
      A[i * n + i] -= cij * cij / dj;
    }
  }
  for (int j = 0; j < n; ++j) {
    real sd = A[j * n + j];
    for (int i = j + 1; i < n; ++i) A[i * n + j] *= sd;
    for (int i = 0; i < j; ++i) A[i * n + j] = 0;
  }
  memcpy(L, A, sizeof(real) * n * n);
}

EXPORT void orc_psd_fix(const real* A, int n, real* fixed, real* inv) {
  real L[36];
  orc_gmw81(A, n, L);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      real acc = 0;
      for (int k = 0; k <= j; ++k) acc += L[i * n + k] * L[j * n + k];
      fixed[i * n + j] = acc;
      fixed[j * n + i] = acc;
    }
  for (int k = 0; k < n; ++k) {
    L[k * n + k] = 1 / L[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      real acc = 0;
      for (int j = k; j < i; ++j) acc += L[i * n + j] * L[j * n + k];
      L[i * n + k] = -acc / L[i * n + i];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      real acc = 0;
      for (int k = i; k < n; ++k) acc += L[k * n + i] * L[k * n + j];
      inv[i * n + j] = acc;
      inv[j * n + i] = acc;
    }
}

/* ------------------------------------------------------------------ */
/* elastic energies in stretch space (refinement/elastic_energy.py)     */
/* ------------------------------------------------------------------ */
static real frob2_s(const real* s) {
  return s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
}
static real det_s(const real* s) {
  return (s[0] * s[1] * s[2] + 2 * s[3] * s[4] * s[5]) - (s[1] * s[4] * s[4] + s[0] * s[5] * s[5] + s[2] * s[3] * s[3]);
}
static void grad_det_s(const real* s, real* g) {
  g[0] = s[1] * s[2] - s[5] * s[5];
  g[1] = s[0] * s[2] - s[4] * s[4];
  g[2] = s[0] * s[1] - s[3] * s[3];
  g[3] = 2 * (s[4] * s[5] - s[2] * s[3]);
  g[4] = 2 * (s[3] * s[5] - s[1] * s[4]);
  g[5] = 2 * (s[3] * s[4] - s[0] * s[5]);
}
static void hess_det_s(const real* s, real* H) {
  const real h[36] = {0, s[2], s[1], 0, 0, -2 * s[5],
                      s[2], 0, s[0], 0, -2 * s[4], 0,
                      s[1], s[0], 0, -2 * s[3], 0, 0,
                      0, 0, -2 * s[3], -2 * s[2], 2 * s[5], 2 * s[4],
                      0, -2 * s[4], 0, 2 * s[5], -2 * s[1], 2 * s[3],
                      -2 * s[5], 0, 0, 2 * s[4], 2 * s[3], -2 * s[0]};
  memcpy(H, h, sizeof h);
}

/* ARAP energy keeps the reference's invariant I2 = s0 s1 + s1 s2 + s2 s0 +
   2 sum off^2 (elastic_energy.py:110-112) while its gradient is that of
   |S|_F^2 (elastic_energy.py:122-124, 147-152). */
EXPORT real orc_arap_energy(const real* s, real mu, real vol) {
  real I1 = s[0] + s[1] + s[2];
  real I2 = s[0] * s[1] + s[1] * s[2] + s[2] * s[0] + 2 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  return vol * (mu * (I2 + 3 - 2 * I1) / 2);
}
EXPORT void orc_arap(const real* s, real mu, real vol, real* E, real* g, real* H, real* Hinv) {
  if (E) *E = orc_arap_energy(s, mu, vol);
  for (int k = 0; k < 6; ++k) g[k] = vol * mu * (2 * SYMD[k] * s[k] - (k < 3 ? 2 : 0)) / 2;
  if (H) {
    memset(H, 0, sizeof(real) * 36);
    memset(Hinv, 0, sizeof(real) * 36);
    for (int k = 0; k < 6; ++k) {
      H[k * 7] = mu * vol * SYMD[k];
      Hinv[k * 7] = (1 / (mu * vol)) * (k < 3 ? (real)1 : (real)0.5);
    }
  }
}
EXPORT real orc_nh_energy(const real* s, real mu, real lam, real vol) {
  real I2 = frob2_s(s), J = det_s(s);
  return vol * (mu * (I2 - 3) / 2 - mu * (J - 1) + lam / 2 * (J - 1) * (J - 1));
}
EXPORT void orc_neohookean(const real* s, real mu, real lam, real vol, real* E, real* g, real* H, real* Hinv, real* Hraw) {
  real J = det_s(s), gJ[6];
  grad_det_s(s, gJ);
  if (E) *E = orc_nh_energy(s, mu, lam, vol);
  for (int k = 0; k < 6; ++k) g[k] = vol * (mu * SYMD[k] * s[k] + (lam * (J - 1) - mu) * gJ[k]);
  if (H) {
    real hJ[36], h[36];
    hess_det_s(s, hJ);
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b)
        h[a * 6 + b] = vol * ((a == b ? mu * SYMD[a] : 0) + lam * gJ[a] * gJ[b] + (lam * (J - 1) - mu) * hJ[a * 6 + b]);
    if (Hraw) memcpy(Hraw, h, sizeof h);
    orc_psd_fix(h, 6, H, Hinv);
  }
}

/* ------------------------------------------------------------------ */
/* contact primitives (refinement/contact.py:16-59, ipc/distance.py)    */
/* ------------------------------------------------------------------ */
EXPORT void orc_barrier(real d, real d0, real d1, const real* q, real* out) {
  real e = 0, de = 0, d2e = 0;
  if (d < d1 && d > d0) {
    real l = (real)log((double)(d / d1));
    e = -(d - d1) * (d - d1) * l;
    de = -(2 * (d - d1) * l + (d - d1) * (d - d1) / d);
    d2e = -(2 * l + 4 * (d - d1) / d - (d - d1) * (d - d1) / (d * d));
  } else if (d <= d0) {
    e = q[0] + q[1] * (d - d0) + (real)0.5 * q[2] * (d - d0) * (d - d0);
    de = q[1] + q[2] * (d - d0);
    d2e = q[2];
  }
  out[0] = e; out[1] = de; out[2] = d2e;
}
static real friction_f0(real y, real ev) {
  if (y >= ev) return y;
  return -y * y * y / (3 * ev * ev) + y * y / ev + ev / 3;
}
static real friction_f1_over_y(real y, real ev) {
  if (y >= ev) return 1 / (y > (real)1e-12 ? y : (real)1e-12);
  return -y / (ev * ev) + 2 / ev;
}
/* lagged IPC friction of one vertex (contact.py:304-387): tangential slip
   u_T = (I - n n^T)(x - x_start), E = mu lamn f0(|u_T|, ev), grad =
   mu lamn (f1/y) u_T, hess = mu lamn (f1/y) (I - n n^T); ev = eps_v dt.
   grad / hess may be null. */
EXPORT real orc_friction(const real* x, const real* xs, const real* n, real lamn, real mu, real ev, real* grad,
                         real* hess) {
  real u[3] = {x[0] - xs[0], x[1] - xs[1], x[2] - xs[2]};
  real un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
  real ut[3] = {u[0] - un * n[0], u[1] - un * n[1], u[2] - un * n[2]};
  real y = (real)sqrt((double)(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2])), sc = mu * lamn;
  real f1y = friction_f1_over_y(y, ev);
  if (grad) for (int k = 0; k < 3; ++k) grad[k] = sc * f1y * ut[k];
  if (hess)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) hess[a * 3 + b] = sc * f1y * ((a == b ? 1 : 0) - n[a] * n[b]);
  return sc * friction_f0(y, ev);
}
/* quaternion (x, y, z, w) -> row-major rotation */
static void quat_to_mat(const real* q, real* R) {
  real x = q[0], y = q[1], z = q[2], w = q[3];
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w); R[2] = 2 * (x * z + y * w);
  R[3] = 2 * (x * y + z * w); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
  R[6] = 2 * (x * z - y * w); R[7] = 2 * (y * z + x * w); R[8] = 1 - 2 * (x * x + y * y);
}
enum { SHAPE_PLANE = 0, SHAPE_CAPSULE = 1 };

/* signed distance of world point x to shape (xf = p(3), q(4) world transform) */
EXPORT real orc_shape_sdf(int type, const real* xf, const real* scale, const real* x, real* grad, real* hess) {
  real R[9];
  quat_to_mat(xf + 3, R);
  real dxw[3] = {x[0] - xf[0], x[1] - xf[1], x[2] - xf[2]};
  real xl[3];
  for (int i = 0; i < 3; ++i) xl[i] = R[0 * 3 + i] * dxw[0] + R[1 * 3 + i] * dxw[1] + R[2 * 3 + i] * dxw[2];
  real d, n[3], h[9];
  memset(h, 0, sizeof h);
  if (type == SHAPE_PLANE) {
    d = xl[2];
    n[0] = 0; n[1] = 0; n[2] = 1;
  } else {
    real r = scale[0], hh = scale[1];
    real az = xl[2] < -hh ? -hh : (xl[2] > hh ? hh : xl[2]);
    real df[3] = {xl[0], xl[1], xl[2] - az};
    real rho = norm3(df);
    real eps = (real)1e-8;
    real rs = rho > eps ? rho : eps;
    n[0] = 0; n[1] = 0; n[2] = 1;
    if (rho > eps) { n[0] = df[0] / rho; n[1] = df[1] / rho; n[2] = df[2] / rho; }
    d = rho - r;
    int cap = (xl[2] > hh || xl[2] < -hh);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        h[i * 3 + j] = ((i == j ? 1 : 0) - (cap ? 0 : (i == 2 && j == 2 ? 1 : 0)) - n[i] * n[j]) / rs;
  }
  if (grad)
    for (int i = 0; i < 3; ++i) grad[i] = R[i * 3 + 0] * n[0] + R[i * 3 + 1] * n[1] + R[i * 3 + 2] * n[2];
  if (hess) {
    real RH[9];
    mm(R, h, RH, 3, 3, 3);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        real acc = 0;
        for (int k = 0; k < 3; ++k) acc += RH[i * 3 + k] * R[j * 3 + k];
        hess[i * 3 + j] = acc;
      }
  }
  return d;
}

/* closest points between segments p1q1 and p2q2 (Ericson 5.1.9), returns the
   parameter along the first segment (collision.py:276-322). */
static real seg_seg(const real* p1, const real* q1, const real* p2, const real* q2, real* c1, real* c2) {
  real d1[3] = {q1[0] - p1[0], q1[1] - p1[1], q1[2] - p1[2]};
  real d2[3] = {q2[0] - p2[0], q2[1] - p2[1], q2[2] - p2[2]};
  real r[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  real a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
  real s, t;
  const real eps = (real)1e-12;
  if (a <= eps && e <= eps) { s = 0; t = 0; }
  else if (a <= eps) { s = 0; t = f / e; t = t < 0 ? 0 : (t > 1 ? 1 : t); }
  else {
    real c = dot3(d1, r);
    if (e <= eps) { t = 0; s = -c / a; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
    else {
      real b = dot3(d1, d2), den = a * e - b * b;
      if (den != 0) { s = (b * f - c * e) / den; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
      else s = 0;
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = -c / a; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
      else if (t > 1) { t = 1; s = (b - c) / a; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
    }
  }
  for (int i = 0; i < 3; ++i) { c1[i] = p1[i] + d1[i] * s; c2[i] = p2[i] + d2[i] * t; }
  return s;
}

/* barycentric weights of the closest point of tri abc to x (Ericson 5.1.5) */
EXPORT void orc_closest_point_triangle_bary(const real* x, const real* a, const real* b, const real* c, real* w) {
  real ab[3], ac[3], ax[3], bx[3], cx[3];
  for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ax[i] = x[i] - a[i]; bx[i] = x[i] - b[i]; cx[i] = x[i] - c[i]; }
  real d1 = dot3(ab, ax), d2 = dot3(ac, ax);
  if (d1 <= 0 && d2 <= 0) { w[0] = 1; w[1] = 0; w[2] = 0; return; }
  real d3 = dot3(ab, bx), d4 = dot3(ac, bx);
  if (d3 >= 0 && d4 <= d3) { w[0] = 0; w[1] = 1; w[2] = 0; return; }
  real d5 = dot3(ab, cx), d6 = dot3(ac, cx);
  if (d6 >= 0 && d5 <= d6) { w[0] = 0; w[1] = 0; w[2] = 1; return; }
  real vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) { real v = d1 / (d1 - d3); w[0] = 1 - v; w[1] = v; w[2] = 0; return; }
  real vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) { real ww = d2 / (d2 - d6); w[0] = 1 - ww; w[1] = 0; w[2] = ww; return; }
  real va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    real ww = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    w[0] = 0; w[1] = 1 - ww; w[2] = ww; return;
  }
  real tot = va + vb + vc;
  if (tot <= 0) { w[0] = w[1] = w[2] = (real)(1.0 / 3.0); return; }
  real v = vb / tot, ww = vc / tot;
  w[0] = 1 - v - ww; w[1] = v; w[2] = ww;
}

static int seg_tri_intersect(const real* p0, const real* p1, const real* a, const real* b, const real* c, real* w) {
  real dir[3], e1[3], e2[3], pv[3], tv[3], qv[3];
  for (int i = 0; i < 3; ++i) { dir[i] = p1[i] - p0[i]; e1[i] = b[i] - a[i]; e2[i] = c[i] - a[i]; tv[i] = p0[i] - a[i]; }
  cross3(dir, e2, pv);
  real det = dot3(e1, pv);
  real scale = norm3(dir) * norm3(e1) * norm3(e2);
  if ((real)fabs((double)det) <= (real)1e-7 * scale + (real)1e-30) return 0;
  real inv = 1 / det;
  real u = dot3(tv, pv) * inv;
  if (u < 0 || u > 1) return 0;
  cross3(tv, e1, qv);
  real v = dot3(dir, qv) * inv;
  if (v < 0 || u + v > 1) return 0;
  real t = dot3(e2, qv) * inv;
  if (t < 0 || t > 1) return 0;
  w[0] = 1 - u - v; w[1] = u; w[2] = v;
  return 1;
}

/* closest tri point to the segment: five candidates + pierce test
   (contact.py:496-550). Returns distance; bary and segment point out. */
EXPORT real orc_closest_point_triangle_segment(const real* a, const real* b, const real* c, const real* p0, const real* p1, real* bary, real* q, int* hit) {
  real w[3], p[3], best_d;
  orc_closest_point_triangle_bary(p0, a, b, c, bary);
  for (int i = 0; i < 3; ++i) { p[i] = bary[0] * a[i] + bary[1] * b[i] + bary[2] * c[i] - p0[i]; q[i] = p0[i]; }
  best_d = norm3(p);
  orc_closest_point_triangle_bary(p1, a, b, c, w);
  for (int i = 0; i < 3; ++i) p[i] = w[0] * a[i] + w[1] * b[i] + w[2] * c[i] - p1[i];
  real d = norm3(p);
  if (d < best_d) { best_d = d; memcpy(bary, w, sizeof w); memcpy(q, p1, 3 * sizeof(real)); }
  const real* E0[3] = {a, b, c};
  const real* E1[3] = {b, c, a};
  for (int e = 0; e < 3; ++e) {
    real c1[3], c2[3];
    real s = seg_seg(E0[e], E1[e], p0, p1, c1, c2);
    real df[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
    d = norm3(df);
    if (d < best_d) {
      best_d = d;
      if (e == 0) { bary[0] = 1 - s; bary[1] = s; bary[2] = 0; }
      else if (e == 1) { bary[0] = 0; bary[1] = 1 - s; bary[2] = s; }
      else { bary[0] = s; bary[1] = 0; bary[2] = 1 - s; }
      memcpy(q, c2, 3 * sizeof(real));
    }
  }
  *hit = seg_tri_intersect(p0, p1, a, b, c, w);
  if (*hit) {
    best_d = 0;
    memcpy(bary, w, sizeof w);
    for (int i = 0; i < 3; ++i) q[i] = w[0] * a[i] + w[1] * b[i] + w[2] * c[i];
  }
  return best_d;
}

/* ------------------------------------------------------------------ */
/* membrane shell over the surface triangles                            */
/* (src/mfem/refinement/shell_energy.py:37-233)                         */
/* ------------------------------------------------------------------ */
/* Inverse 2D rest-shape matrix W (row-major 2x2) of rest triangle r0 r1 r2
   in the in-plane frame u = e1/|e1|, v = n x u, n the unit normal; returns
   the rest area, or 0 with W = 0 for a degenerate triangle (|e1 x e2| <
   1e-12) (shell_energy.py:37-63). */
EXPORT real orc_tri_rest_frame(const real* r0, const real* r1, const real* r2, real* W) {
  real e1[3], e2[3], n[3];
  for (int k = 0; k < 3; ++k) { e1[k] = r1[k] - r0[k]; e2[k] = r2[k] - r0[k]; }
  cross3(e1, e2, n);
  real ta = norm3(n);
  if (ta < (real)1e-12) { W[0] = W[1] = W[2] = W[3] = 0; return 0; }
  real u[3], nn[3], v[3], le = norm3(e1);
  for (int k = 0; k < 3; ++k) { u[k] = e1[k] / le; nn[k] = n[k] / ta; }
  cross3(nn, u, v);
  real a = dot3(e1, u), b = dot3(e2, u), c = dot3(e1, v), d = dot3(e2, v);
  real det = a * d - b * c;
  W[0] = d / det; W[1] = -b / det; W[2] = -c / det; W[3] = a / det;
  return (real)0.5 * ta;
}

/* Membrane energy mu A |F - R(F)|^2 / 2 of one triangle, F = [x1 - x0,
   x2 - x0] W (3x2), R its polar factor, through the closed forms of the
   2x2 C = F^T F: sigma1 + sigma2 = sqrt(tr C + 2 sqrt(det C)) and
   F - R = F (I - S^-1), S = (C + sqrt(det C) I) / (sigma1 + sigma2); zero
   stress when sqrt(det C) < 1e-10 (shell_energy.py:66-125, 190-229).
   grad (9, may be null) = dE/dx0, dE/dx1, dE/dx2.  Returns the energy. */
EXPORT real orc_membrane(const real* x0, const real* x1, const real* x2, const real* W, real area, real mu,
                         real* grad) {
  real d1[3], d2[3], f1[3], f2[3];
  for (int k = 0; k < 3; ++k) {
    d1[k] = x1[k] - x0[k];
    d2[k] = x2[k] - x0[k];
    f1[k] = d1[k] * W[0] + d2[k] * W[2];
    f2[k] = d1[k] * W[1] + d2[k] * W[3];
  }
  real c11 = dot3(f1, f1), c12 = dot3(f1, f2), c22 = dot3(f2, f2);
  real dc = c11 * c22 - c12 * c12;
  if (dc < 0) dc = 0;
  real tr2 = c11 + c22 + 2 * (real)sqrt((double)dc);
  real ssum = (real)sqrt((double)(tr2 > (real)1e-20 ? tr2 : (real)1e-20));
  real E = mu * area * (real)0.5 * (c11 + c22 - 2 * ssum + 2);
  if (grad) {
    real s = (real)sqrt((double)dc), p1[3] = {0, 0, 0}, p2[3] = {0, 0, 0};
    if (s >= (real)1e-10) {
      real den = ssum * s;
      real i11 = (c22 + s) / den, i12 = -c12 / den, i22 = (c11 + s) / den;
      for (int k = 0; k < 3; ++k) {
        p1[k] = mu * area * (f1[k] - (f1[k] * i11 + f2[k] * i12));
        p2[k] = mu * area * (f2[k] - (f1[k] * i12 + f2[k] * i22));
      }
    }
    for (int k = 0; k < 3; ++k) {
      real g1 = p1[k] * W[0] + p2[k] * W[1];
      real g2 = p1[k] * W[2] + p2[k] * W[3];
      grad[3 + k] = g1;
      grad[6 + k] = g2;
      grad[k] = -(g1 + g2);
    }
  }
  return E;
}

/* ------------------------------------------------------------------ */
/* solver context                                                       */
/* ------------------------------------------------------------------ */
typedef struct {
  int nv, nt, ntris, nshapes, npins;
  /* model */
  int* tets; real* dminv; real* mat; real* mass; real* inv_mass; real* vol;
  int* pin_idx; real* pin_pos;
  int* tris;
  int* shape_type; real* shape_xf; real* shape_scale; real* shape_mu;
  /* params */
  real p[32];
  /* pattern */
  int* rowptr; int* col; int nnzb; int* diag; int* tet_slot;
  real* val;
  /* state */
  real *x, *v, *s, *lam, *xt, *xstart, *xin;
  real *lag_n, *lag_fn; int* lag_sid;
  /* per tet */
  real *c, *G, *gs, *Hs, *Hsinv, *Hl, *gl, *psi;
  /* per vertex */
  real *gx, *gB, *gD, *HB, *HD, *dist, *ddist, *hdist; int* sid;
  real *eB, *eD;
  real *g, *dx, *ds, *Minv;
  /* per tri (capsule contact) */
  real *tdist, *tbary, *tnorm; int* tsid;
  /* scratch */
  real *r, *z, *pp, *q, *xtrial, *strial;
  /* stats: per Newton iteration */
  int stat_n; real stats[4096];
  /* self-contact: surface vertices and edges; barrier gradient and
     diagonal Hessian blocks (FP64, sym6 order 00 11 22 01 02 12) */
  int npts, nedges; int* pts; int* edges;
  double *sc_g, *sc_h;
  /* membrane shell: rest positions, per-triangle inverse rest-shape matrix
     and rest area (set by orc_shell_refresh), energy per triangle,
     gradient per vertex */
  real *rest, *tW, *tarea, *eS, *gS;
} orc_ctx;

enum {
  P_ENERGY = 0, P_GX, P_GY, P_GZ, P_ATTACH_K, P_D0, P_D1, P_KAPPA, P_EPSV,
  P_ITERS, P_CG_MAXIT, P_CG_TOL, P_PRECOND, P_LS, P_LS_MAX, P_LS_A0, P_LS_TAU,
  P_LS_C, P_LS_THR, P_SING, P_TRI_CONTACT, P_GUESS, P_SC, P_SC_DHAT, P_SC_KAPPA, P_SC_S, P_SHELL_MU,
  P_QB0, P_QB1, P_QB2, P_NPARAMS
};
static const char* PNAMES[] = {"energy", "gx", "gy", "gz", "attach_k", "d0", "d1", "kappa", "eps_v",
                               "iterations", "cg_maxiter", "cg_tol", "precond", "line_search", "ls_max", "ls_alpha0",
                               "ls_tau", "ls_c", "ls_thr", "sing_thr", "tri_contact", "guess", "self_contact",
                               "sc_dhat", "sc_kappa", "sc_ccd_s", "shell_mu", NULL};

#define ALLOC(ptr, n) ptr = calloc((size_t)((n) > 0 ? (n) : 1), sizeof(*(ptr)))

EXPORT orc_ctx* orc_create(int nv, int nt, int ntris, int nshapes, int npins) {
  orc_ctx* C = calloc(1, sizeof(orc_ctx));
  C->nv = nv; C->nt = nt; C->ntris = ntris; C->nshapes = nshapes; C->npins = npins;
  ALLOC(C->tets, 4 * nt); ALLOC(C->dminv, 9 * nt); ALLOC(C->mat, 3 * nt); ALLOC(C->vol, nt);
  ALLOC(C->mass, nv); ALLOC(C->inv_mass, nv);
  ALLOC(C->pin_idx, npins); ALLOC(C->pin_pos, 3 * npins);
  ALLOC(C->tris, 3 * ntris);
  ALLOC(C->shape_type, nshapes); ALLOC(C->shape_xf, 7 * nshapes); ALLOC(C->shape_scale, 3 * nshapes); ALLOC(C->shape_mu, nshapes);
  ALLOC(C->x, 3 * nv); ALLOC(C->v, 3 * nv); ALLOC(C->s, 6 * nt); ALLOC(C->lam, 6 * nt);
  ALLOC(C->xt, 3 * nv); ALLOC(C->xstart, 3 * nv); ALLOC(C->xin, 3 * nv);
  ALLOC(C->lag_n, 3 * nv); ALLOC(C->lag_fn, nv); ALLOC(C->lag_sid, nv);
  ALLOC(C->c, 6 * nt); ALLOC(C->G, 72 * nt); ALLOC(C->gs, 6 * nt); ALLOC(C->Hs, 36 * nt); ALLOC(C->Hsinv, 36 * nt);
  ALLOC(C->Hl, 36 * nt); ALLOC(C->gl, 6 * nt); ALLOC(C->psi, nt);
  ALLOC(C->gx, 3 * nv); ALLOC(C->gB, 3 * nv); ALLOC(C->gD, 3 * nv); ALLOC(C->HB, 9 * nv); ALLOC(C->HD, 9 * nv);
  ALLOC(C->dist, nv); ALLOC(C->ddist, 3 * nv); ALLOC(C->hdist, 9 * nv); ALLOC(C->sid, nv);
  ALLOC(C->eB, nv); ALLOC(C->eD, nv);
  ALLOC(C->g, 3 * nv); ALLOC(C->dx, 3 * nv); ALLOC(C->ds, 6 * nt); ALLOC(C->Minv, 9 * nv);
  ALLOC(C->tdist, ntris); ALLOC(C->tbary, 3 * ntris); ALLOC(C->tnorm, 3 * ntris); ALLOC(C->tsid, ntris);
  ALLOC(C->r, 3 * nv); ALLOC(C->z, 3 * nv); ALLOC(C->pp, 3 * nv); ALLOC(C->q, 3 * nv);
  ALLOC(C->xtrial, 3 * nv); ALLOC(C->strial, 6 * nt);
  ALLOC(C->tet_slot, 16 * nt); ALLOC(C->diag, nv); ALLOC(C->rowptr, nv + 1);
  /* defaults: RefinementSolver kwargs (refinement/solver.py:56-128, 241-245) */
  C->p[P_ENERGY] = 0; C->p[P_GZ] = (real)-9.81; C->p[P_ATTACH_K] = 10;
  C->p[P_D0] = (real)1e-2; C->p[P_D1] = (real)1e-1; C->p[P_KAPPA] = (real)1e7; C->p[P_EPSV] = (real)1e-3;
  C->p[P_ITERS] = 8; C->p[P_CG_MAXIT] = 1000; C->p[P_CG_TOL] = (real)1e-6; C->p[P_PRECOND] = 1;
  C->p[P_LS] = 1; C->p[P_LS_MAX] = 20; C->p[P_LS_A0] = 1; C->p[P_LS_TAU] = (real)0.5; C->p[P_LS_C] = (real)0.01;
  C->p[P_LS_THR] = (real)1e-8; C->p[P_SING] = (real)1e-20; C->p[P_TRI_CONTACT] = 1;
  C->p[P_SC] = 0; C->p[P_SC_DHAT] = (real)1e-3; C->p[P_SC_KAPPA] = (real)1e7; C->p[P_SC_S] = (real)0.1;
  ALLOC(C->sc_g, 3 * nv); ALLOC(C->sc_h, 6 * nv);
  ALLOC(C->rest, 3 * nv); ALLOC(C->tW, 4 * ntris); ALLOC(C->tarea, ntris); ALLOC(C->eS, ntris); ALLOC(C->gS, 3 * nv);
  return C;
}

/* surface vertices and edges for self-contact (the triangles are tris[]) */
EXPORT void orc_set_surface(orc_ctx* C, int npts, const int* pts, int nedges, const int* edges) {
  free(C->pts); free(C->edges);
  C->npts = npts; C->nedges = nedges;
  ALLOC(C->pts, npts); ALLOC(C->edges, 2 * nedges);
  memcpy(C->pts, pts, sizeof(int) * (size_t)npts);
  memcpy(C->edges, edges, sizeof(int) * 2 * (size_t)nedges);
}

EXPORT int orc_set_param(orc_ctx* C, const char* name, double value) {
  for (int i = 0; PNAMES[i]; ++i)
    if (!strcmp(PNAMES[i], name)) { C->p[i] = (real)value; return 0; }
  return -1;
}
EXPORT double orc_get_param(orc_ctx* C, const char* name) {
  for (int i = 0; PNAMES[i]; ++i)
    if (!strcmp(PNAMES[i], name)) return (double)C->p[i];
  return NAN;
}

/* buffer table for the ctypes wrapper: name -> pointer, element count, kind (0 real, 1 int, 2 double) */
EXPORT void* orc_buf(orc_ctx* C, const char* name, int64_t* count, int* is_int) {
#define B(nm, ptr, n, k) if (!strcmp(name, nm)) { *count = (int64_t)(n); *is_int = (k); return (void*)(ptr); }
  B("tets", C->tets, 4 * C->nt, 1) B("dminv", C->dminv, 9 * C->nt, 0) B("materials", C->mat, 3 * C->nt, 0)
  B("vol", C->vol, C->nt, 0) B("mass", C->mass, C->nv, 0) B("inv_mass", C->inv_mass, C->nv, 0)
  B("pin_idx", C->pin_idx, C->npins, 1) B("pin_pos", C->pin_pos, 3 * C->npins, 0) B("tris", C->tris, 3 * C->ntris, 1)
  B("shape_type", C->shape_type, C->nshapes, 1) B("shape_xf", C->shape_xf, 7 * C->nshapes, 0)
  B("shape_scale", C->shape_scale, 3 * C->nshapes, 0) B("shape_mu", C->shape_mu, C->nshapes, 0)
  B("x", C->x, 3 * C->nv, 0) B("v", C->v, 3 * C->nv, 0) B("s", C->s, 6 * C->nt, 0) B("lam", C->lam, 6 * C->nt, 0)
  B("x_tilde", C->xt, 3 * C->nv, 0) B("x_start", C->xstart, 3 * C->nv, 0)
  B("lag_n", C->lag_n, 3 * C->nv, 0) B("lag_fn", C->lag_fn, C->nv, 0) B("lag_sid", C->lag_sid, C->nv, 1)
  B("c", C->c, 6 * C->nt, 0) B("G", C->G, 72 * C->nt, 0) B("gs", C->gs, 6 * C->nt, 0) B("Hs", C->Hs, 36 * C->nt, 0)
  B("Hsinv", C->Hsinv, 36 * C->nt, 0) B("Hl", C->Hl, 36 * C->nt, 0) B("gl", C->gl, 6 * C->nt, 0) B("psi", C->psi, C->nt, 0)
  B("gx", C->gx, 3 * C->nv, 0) B("gB", C->gB, 3 * C->nv, 0) B("gD", C->gD, 3 * C->nv, 0) B("HB", C->HB, 9 * C->nv, 0)
  B("HD", C->HD, 9 * C->nv, 0) B("dist", C->dist, C->nv, 0) B("eB", C->eB, C->nv, 0) B("eD", C->eD, C->nv, 0)
  B("g", C->g, 3 * C->nv, 0) B("dx", C->dx, 3 * C->nv, 0) B("ds", C->ds, 6 * C->nt, 0) B("Minv", C->Minv, 9 * C->nv, 0)
  B("rowptr", C->rowptr, C->nv + 1, 1) B("col", C->col, C->nnzb, 1) B("val", C->val, 9 * C->nnzb, 0)
  B("diag", C->diag, C->nv, 1) B("tdist", C->tdist, C->ntris, 0) B("tbary", C->tbary, 3 * C->ntris, 0)
  B("stats", C->stats, 4096, 0) B("sc_g", C->sc_g, 3 * C->nv, 2) B("sc_h", C->sc_h, 6 * C->nv, 2)
  B("rest", C->rest, 3 * C->nv, 0) B("tW", C->tW, 4 * C->ntris, 0) B("tarea", C->tarea, C->ntris, 0)
  B("eS", C->eS, C->ntris, 0) B("gS", C->gS, 3 * C->nv, 0)
#undef B
  *count = 0; *is_int = 0;
  return NULL;
}

static int cmp_int(const void* a, const void* b) { return (*(const int*)a > *(const int*)b) - (*(const int*)a < *(const int*)b); }

/* Builds the vertex-adjacency CSR pattern (diagonal included, columns sorted)
   and each tet's 16 block destinations; computes signed rest volumes. */
EXPORT int orc_finalize(orc_ctx* C) {
  int nv = C->nv, nt = C->nt;
  int* deg = calloc(nv, sizeof(int));
  for (int e = 0; e < nt; ++e)
    for (int a = 0; a < 4; ++a) {
      int i = C->tets[4 * e + a];
      if (i < 0 || i >= nv) { free(deg); return -1; }
      deg[i] += 4;
    }
  int* off = calloc(nv + 1, sizeof(int));
  for (int i = 0; i < nv; ++i) off[i + 1] = off[i] + deg[i] + 1;
  int* lst = malloc(sizeof(int) * (off[nv] > 0 ? off[nv] : 1));
  int* fill = calloc(nv, sizeof(int));
  for (int i = 0; i < nv; ++i) lst[off[i] + fill[i]++] = i;
  for (int e = 0; e < nt; ++e)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        int i = C->tets[4 * e + a], j = C->tets[4 * e + b];
        if (i != j) lst[off[i] + fill[i]++] = j;
      }
  C->rowptr[0] = 0;
  int total = 0;
  for (int i = 0; i < nv; ++i) {
    qsort(lst + off[i], fill[i], sizeof(int), cmp_int);
    int u = 0;
    for (int k = 0; k < fill[i]; ++k)
      if (k == 0 || lst[off[i] + k] != lst[off[i] + k - 1]) lst[off[i] + u++] = lst[off[i] + k];
    fill[i] = u;
    total += u;
    C->rowptr[i + 1] = total;
  }
  free(C->col); free(C->val);
  C->nnzb = total;
  C->col = malloc(sizeof(int) * (total > 0 ? total : 1));
  C->val = calloc((size_t)9 * (total > 0 ? total : 1), sizeof(real));
  for (int i = 0; i < nv; ++i) {
    memcpy(C->col + C->rowptr[i], lst + off[i], sizeof(int) * fill[i]);
    for (int k = C->rowptr[i]; k < C->rowptr[i + 1]; ++k)
      if (C->col[k] == i) C->diag[i] = k;
  }
  for (int e = 0; e < nt; ++e)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        int i = C->tets[4 * e + a], j = C->tets[4 * e + b];
        int lo = C->rowptr[i], hi = C->rowptr[i + 1] - 1, pos = -1;
        while (lo <= hi) {
          int mid = (lo + hi) / 2;
          if (C->col[mid] == j) { pos = mid; break; }
          if (C->col[mid] < j) lo = mid + 1; else hi = mid - 1;
        }
        C->tet_slot[16 * e + 4 * a + b] = pos;
      }
  for (int e = 0; e < nt; ++e) C->vol[e] = 1 / (6 * det3(C->dminv + 9 * e));
  free(deg); free(off); free(lst); free(fill);
  real qb[3];
  orc_barrier(C->p[P_D0], 0, C->p[P_D1], qb /* unused q for log branch */, qb);
  C->p[P_QB0] = qb[0]; C->p[P_QB1] = qb[1]; C->p[P_QB2] = qb[2];
  return 0;
}

/* rest stretch s = sym6(S(F)) at the given positions (refinement/kernels.py:55-75) */
EXPORT void orc_init_stretch(orc_ctx* C) {
  for (int e = 0; e < C->nt; ++e) {
    real xe[12], c[6], zero[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 4; ++a)
      for (int k = 0; k < 3; ++k) xe[3 * a + k] = C->x[3 * C->tets[4 * e + a] + k];
    orc_tet_constraints(xe, C->dminv + 9 * e, zero, c, NULL, C->s + 6 * e);
  }
}

/* ---- contact at positions X (refinement/contact.py:98-258, 304-387) ---- */
static void vertex_distance(orc_ctx* C, const real* x, real* d, real* grad, real* hess, int* sid) {
  real dm = (real)1e8, gm[3] = {0, 0, 0}, hm[9] = {0};
  int im = -1;
  for (int k = 0; k < C->nshapes; ++k) {
    real gr[3], hs[9];
    real dd = orc_shape_sdf(C->shape_type[k], C->shape_xf + 7 * k, C->shape_scale + 3 * k, x, gr, hs);
    if (dd < dm) { dm = dd; memcpy(gm, gr, sizeof gr); memcpy(hm, hs, sizeof hs); im = k; }
  }
  *d = dm;
  if (grad) memcpy(grad, gm, sizeof gm);
  if (hess) memcpy(hess, hm, sizeof hm);
  if (sid) *sid = im;
}

static void tri_capsule_shapes(int nshapes, const int* shape_type, const real* shape_xf, const real* shape_scale,
                               const real* a, const real* b, const real* c, real* dmin, real* bary, real* nrm,
                               int* sid) {
  *dmin = (real)1e8;
  bary[0] = bary[1] = bary[2] = (real)(1.0 / 3.0);
  nrm[0] = nrm[1] = nrm[2] = 0;
  *sid = -1;
  for (int k = 0; k < nshapes; ++k) {
    if (shape_type[k] != SHAPE_CAPSULE) continue;
    const real* xf = shape_xf + 7 * k;
    real R[9];
    quat_to_mat(xf + 3, R);
    real hh = shape_scale[3 * k + 1], rad = shape_scale[3 * k + 0];
    real p0[3], p1[3];
    for (int i = 0; i < 3; ++i) { p0[i] = xf[i] - R[i * 3 + 2] * hh; p1[i] = xf[i] + R[i * 3 + 2] * hh; }
    real w[3], q[3];
    int hit;
    real dist = orc_closest_point_triangle_segment(a, b, c, p0, p1, w, q, &hit);
    real d = dist - rad;
    if (d < *dmin) {
      real p[3], n[3];
      for (int i = 0; i < 3; ++i) p[i] = w[0] * a[i] + w[1] * b[i] + w[2] * c[i];
      if (!hit && dist > (real)1e-8) {
        for (int i = 0; i < 3; ++i) n[i] = (p[i] - q[i]) / dist;
      } else {
        real e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        cross3(e1, e2, n);
        real nl = norm3(n);
        if (nl > (real)1e-20) { n[0] /= nl; n[1] /= nl; n[2] /= nl; }
        else { n[0] = 0; n[1] = 0; n[2] = 1; }
        real ax[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        real pp0[3] = {p[0] - p0[0], p[1] - p0[1], p[2] - p0[2]};
        real aa = dot3(ax, ax);
        real t = dot3(pp0, ax) / (aa > (real)1e-20 ? aa : (real)1e-20);
        real sgn = t < (real)0.5 ? -1 : 1;
        if (sgn * dot3(n, ax) < 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
      }
      *dmin = d;
      memcpy(bary, w, 3 * sizeof(real));
      memcpy(nrm, n, 3 * sizeof(real));
      *sid = k;
    }
  }
}

static void tri_capsule(orc_ctx* C, const real* a, const real* b, const real* c, real* dmin, real* bary, real* nrm,
                        int* sid) {
  tri_capsule_shapes(C->nshapes, C->shape_type, C->shape_xf, C->shape_scale, a, b, c, dmin, bary, nrm, sid);
}

/* One surface triangle against capsules (contact.py:553-775): distance d of
   the closest triangle point to the nearest capsule minus its radius, the
   barycentric weights w and normal n of that point, energy kappa b(d),
   gradient w_a kappa b'(d) n per corner (grad[3][3]) and the Gauss-Newton
   Hessian blocks w_a w_b kappa max(b''(d), 0) n n^T (hess[3][3][3][3],
   corner a, corner b, row, column); zero outside d < d1.  q = b(d0), b'(d0),
   b''(d0) of the quadratic extension.  Returns the energy. */
EXPORT real orc_tri_barrier(const real* a, const real* b, const real* c, int nshapes, const int* shape_type,
                            const real* shape_xf, const real* shape_scale, real d0, real d1, real kappa,
                            const real* q, real* d_out, real* bary, real* nrm, real* grad, real* hess) {
  int sid;
  real d;
  tri_capsule_shapes(nshapes, shape_type, shape_xf, shape_scale, a, b, c, &d, bary, nrm, &sid);
  *d_out = d;
  real bb[3] = {0, 0, 0};
  if (d < d1) orc_barrier(d, d0, d1, q, bb);
  real h2 = bb[2] > 0 ? bb[2] : 0;
  for (int i = 0; i < 3; ++i)
    for (int r = 0; r < 3; ++r) grad[3 * i + r] = d < d1 ? bary[i] * kappa * bb[1] * nrm[r] : 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc)
          hess[((i * 3 + j) * 3 + r) * 3 + cc] = d < d1 ? bary[i] * bary[j] * kappa * h2 * nrm[r] * nrm[cc] : 0;
  return d < d1 ? kappa * bb[0] : 0;
}

/* start of step (solver.py:499-520): x_tilde, friction anchor, lagged force */
EXPORT void orc_begin_step(orc_ctx* C, real dt) {
  real g[3] = {C->p[P_GX], C->p[P_GY], C->p[P_GZ]};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < C->nv; ++i)
    for (int k = 0; k < 3; ++k) {
      C->xin[3 * i + k] = C->x[3 * i + k];
      C->xt[3 * i + k] = C->x[3 * i + k] + C->v[3 * i + k] * dt + g[k] * dt * dt / 2;
      C->xstart[3 * i + k] = C->x[3 * i + k];
    }
  real mumax = 0;
  for (int k = 0; k < C->nshapes; ++k) if (C->shape_mu[k] > mumax) mumax = C->shape_mu[k];
  if (C->nshapes == 0 || mumax <= 0) return;
  const real* qb = C->p + P_QB0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < C->nv; ++i) {
    real d, gr[3];
    int sid;
    vertex_distance(C, C->x + 3 * i, &d, gr, NULL, &sid);
    real lamn = 0, n[3] = {0, 0, 0};
    int s = -1;
    if (d < C->p[P_D1]) {
      real b[3];
      orc_barrier(d, C->p[P_D0], C->p[P_D1], qb, b);
      lamn = C->p[P_KAPPA] * (-b[1] > 0 ? -b[1] : 0);
      memcpy(n, gr, sizeof n);
      s = sid;
    }
    memcpy(C->lag_n + 3 * i, n, sizeof n);
    C->lag_fn[i] = lamn;
    C->lag_sid[i] = s;
  }
}

static void friction_at(orc_ctx* C, int i, const real* x, real dt, real* E, real* grad, real* hess) {
  *E = 0;
  if (grad) memset(grad, 0, 3 * sizeof(real));
  if (hess) memset(hess, 0, 9 * sizeof(real));
  real lamn = C->lag_fn[i];
  int sid = C->lag_sid[i];
  if (!(lamn > 0 && sid >= 0)) return;
  real mu = C->shape_mu[sid];
  if (!(mu > 0)) return;
  *E = orc_friction(x, C->xstart + 3 * i, C->lag_n + 3 * i, lamn, mu, C->p[P_EPSV] * dt, grad, hess);
}

static int pin_of(orc_ctx* C, int i) {
  for (int k = 0; k < C->npins; ++k) if (C->pin_idx[k] == i) return k;
  return -1;
}

/* _update_system (solver.py:633-639, 930-980) */
/* ---------------------------------------------------------------------------
   Self-contact: IPC barrier on point-triangle and edge-edge pairs of surface
   primitives inside the Newton solve.  The reference's live solve has no
   such term (the hook at src/mfem/solver.py:443-447 is commented out); the
   primitives follow src/mfem/collision.py:220-322 (Ericson closest points)
   and the barrier b(d) = -(d - dhat)^2 ln(d / dhat) follows
   collision.py:529-546.  Energy kappa b(d) per pair below dhat, gradient
   kappa b'(d) cw_j n per vertex (n the unit vector between the closest
   points, cw_j the vertex's weight in their difference), Gauss-Newton
   diagonal blocks kappa max(b'', 0) cw_j^2 n n^T; the line search's first
   step is cut to the additive-CCD time of impact (Li, Kaufman, Jiang 2021,
   Alg. 1).  Every pair is enumerated here (boxes inflated by dhat / 2 as a
   prefilter), in FP64 whatever the build's real type.
   --------------------------------------------------------------------------- */
static double dot3d(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void cp_tri_bary_d(const double* x, const double* a, const double* b, const double* c, double* w) {
  double ab[3], ac[3], ax[3], bx[3], cx[3];
  for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ax[i] = x[i] - a[i]; bx[i] = x[i] - b[i]; cx[i] = x[i] - c[i]; }
  double d1 = dot3d(ab, ax), d2 = dot3d(ac, ax);
  if (d1 <= 0 && d2 <= 0) { w[0] = 1; w[1] = 0; w[2] = 0; return; }
  double d3 = dot3d(ab, bx), d4 = dot3d(ac, bx);
  if (d3 >= 0 && d4 <= d3) { w[0] = 0; w[1] = 1; w[2] = 0; return; }
  double d5 = dot3d(ab, cx), d6 = dot3d(ac, cx);
  if (d6 >= 0 && d5 <= d6) { w[0] = 0; w[1] = 0; w[2] = 1; return; }
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) { double v = d1 / (d1 - d3); w[0] = 1 - v; w[1] = v; w[2] = 0; return; }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) { double t = d2 / (d2 - d6); w[0] = 1 - t; w[1] = 0; w[2] = t; return; }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    w[0] = 0; w[1] = 1 - t; w[2] = t; return;
  }
  double tot = va + vb + vc;
  if (tot <= 0) { w[0] = w[1] = w[2] = 1.0 / 3.0; return; }
  double v = vb / tot, t = vc / tot;
  w[0] = 1 - v - t; w[1] = v; w[2] = t;
}

/* parameters (s, t) of the closest points of segments p1q1, p2q2 (Ericson 5.1.9) */
static void seg_seg_d(const double* p1, const double* q1, const double* p2, const double* q2, double* s_out, double* t_out) {
  double d1[3], d2[3], r[3];
  for (int i = 0; i < 3; ++i) { d1[i] = q1[i] - p1[i]; d2[i] = q2[i] - p2[i]; r[i] = p1[i] - p2[i]; }
  double a = dot3d(d1, d1), e = dot3d(d2, d2), f = dot3d(d2, r), eps = 1e-12, s, t;
  if (a <= eps && e <= eps) { s = 0; t = 0; }
  else if (a <= eps) { s = 0; t = f / e; t = t < 0 ? 0 : (t > 1 ? 1 : t); }
  else {
    double c = dot3d(d1, r);
    if (e <= eps) { t = 0; s = -c / a; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
    else {
      double b = dot3d(d1, d2), den = a * e - b * b;
      s = den != 0 ? (b * f - c * e) / den : 0; s = s < 0 ? 0 : (s > 1 ? 1 : s);
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = -c / a; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
      else if (t > 1) { t = 1; s = (b - c) / a; s = s < 0 ? 0 : (s > 1 ? 1 : s); }
    }
  }
  *s_out = s; *t_out = t;
}

/* distance of pair kind (0: point X0 vs triangle X1 X2 X3; 1: edge X0X1 vs
   edge X2X3), the vertices' weights cw and the unit vector n */
EXPORT double orc_sc_distance(int kind, const double* X, double* cw, double* n) {
  double dv[3];
  if (kind == 0) {
    double w[3];
    cp_tri_bary_d(X, X + 3, X + 6, X + 9, w);
    for (int c = 0; c < 3; ++c) dv[c] = X[c] - (w[0] * X[3 + c] + w[1] * X[6 + c] + w[2] * X[9 + c]);
    cw[0] = 1; cw[1] = -w[0]; cw[2] = -w[1]; cw[3] = -w[2];
  } else {
    double s, t;
    seg_seg_d(X, X + 3, X + 6, X + 9, &s, &t);
    for (int c = 0; c < 3; ++c) dv[c] = (X[c] + s * (X[3 + c] - X[c])) - (X[6 + c] + t * (X[9 + c] - X[6 + c]));
    cw[0] = 1 - s; cw[1] = s; cw[2] = -(1 - t); cw[3] = -t;
  }
  double d = sqrt(dot3d(dv, dv)), inv = d > 0 ? 1.0 / d : 0.0;
  for (int c = 0; c < 3; ++c) n[c] = dv[c] * inv;
  return d;
}

static void ipc64(double d, double dh, double* e, double* de, double* d2e) {
  double l = log(d / dh), u = d - dh;
  *e = -u * u * l;
  *de = -(2 * u * l + u * u / d);
  *d2e = -(2 * l + 4 * u / d - u * u / (d * d));
}

/* additive CCD time of impact of pair kind along displacements P (<= 1) */
EXPORT double orc_sc_accd(int kind, const double* X0, const double* P0, double s) {
  double X[12], P[12], m[3] = {0, 0, 0}, len[4], cw[4], n[3];
  memcpy(X, X0, sizeof X);
  memcpy(P, P0, sizeof P);
  for (int j = 0; j < 4; ++j) for (int c = 0; c < 3; ++c) m[c] += 0.25 * P[3 * j + c];
  for (int j = 0; j < 4; ++j) {
    for (int c = 0; c < 3; ++c) P[3 * j + c] -= m[c];
    len[j] = sqrt(dot3d(P + 3 * j, P + 3 * j));
  }
  double lp = kind == 0 ? len[0] + fmax(len[1], fmax(len[2], len[3])) : fmax(len[0], len[1]) + fmax(len[2], len[3]);
  if (!(lp > 0)) return 1.0;
  double d = orc_sc_distance(kind, X, cw, n);
  if (!(d > 0)) return 0.0;
  double g = s * d, t = 0, tl = (1 - s) * d / lp;
  for (int it = 0; it < 256; ++it) {
    for (int k = 0; k < 12; ++k) X[k] += tl * P[k];
    d = orc_sc_distance(kind, X, cw, n);
    if (t > 0 && d < g) return t;
    t += tl;
    if (t > 1) return 1.0;
    tl = 0.9 * d / lp;
  }
  return t;
}

/* box of primitive ids[0..nv-1] at x (and x + dxs when given), inflated by r */
static void sc_box(const real* x, const real* dxs, const int* ids, int nv, double r, double* lo, double* hi) {
  for (int c = 0; c < 3; ++c) { lo[c] = INFINITY; hi[c] = -INFINITY; }
  for (int j = 0; j < nv; ++j)
    for (int c = 0; c < 3; ++c) {
      double a = (double)x[3 * ids[j] + c], b = dxs ? a + (double)dxs[3 * ids[j] + c] : a;
      lo[c] = fmin(lo[c], fmin(a, b));
      hi[c] = fmax(hi[c], fmax(a, b));
    }
  for (int c = 0; c < 3; ++c) { lo[c] -= r; hi[c] += r; }
}

/* mode 0: gradient / diagonal Hessian blocks at x into sc_g / sc_h (returns
   the energy); mode 1: energy at x; mode 2: the smallest time of impact
   along dxs (returns it, <= 1).  Pairs whose boxes (over x, and x + dxs in
   mode 2) inflated by dhat / 2 do not overlap are skipped: their distance is
   at least dhat all along. */
static double sc_pass(orc_ctx* C, const real* x, const real* dxs, int mode) {
  const double dh = C->p[P_SC_DHAT], kap = C->p[P_SC_KAPPA], cs = C->p[P_SC_S], r = 0.5 * dh;
  const real* sw = mode == 2 ? dxs : NULL;
  double E = 0, toi = 1.0;
  if (mode == 0) {
    memset(C->sc_g, 0, sizeof(double) * 3 * C->nv);
    memset(C->sc_h, 0, sizeof(double) * 6 * C->nv);
  }
  /* boxes of the points, triangles and edges */
  double* bp = malloc(sizeof(double) * 6 * (C->npts > 0 ? C->npts : 1));
  double* bt = malloc(sizeof(double) * 6 * (C->ntris > 0 ? C->ntris : 1));
  double* be = malloc(sizeof(double) * 6 * (C->nedges > 0 ? C->nedges : 1));
  for (int a = 0; a < C->npts; ++a) sc_box(x, sw, C->pts + a, 1, r, bp + 6 * a, bp + 6 * a + 3);
  for (int t = 0; t < C->ntris; ++t) sc_box(x, sw, C->tris + 3 * t, 3, r, bt + 6 * t, bt + 6 * t + 3);
  for (int e = 0; e < C->nedges; ++e) sc_box(x, sw, C->edges + 2 * e, 2, r, be + 6 * e, be + 6 * e + 3);
  for (int kind = 0; kind < 2; ++kind) {
    int nA = kind == 0 ? C->npts : C->nedges, nB = kind == 0 ? C->ntris : C->nedges;
    const double* BA = kind == 0 ? bp : be;
    const double* BB = kind == 0 ? bt : be;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : E) reduction(min : toi)
    for (int a = 0; a < nA; ++a)
      for (int bidx = kind == 0 ? 0 : a + 1; bidx < nB; ++bidx) {
        const double* A = BA + 6 * a;
        const double* B = BB + 6 * bidx;
        if (A[0] > B[3] || B[0] > A[3] || A[1] > B[4] || B[1] > A[4] || A[2] > B[5] || B[2] > A[5]) continue;
        int v[4];
        if (kind == 0) {
          const int* t = C->tris + 3 * bidx;
          v[0] = C->pts[a]; v[1] = t[0]; v[2] = t[1]; v[3] = t[2];
          if (v[0] == v[1] || v[0] == v[2] || v[0] == v[3]) continue;
        } else {
          v[0] = C->edges[2 * a]; v[1] = C->edges[2 * a + 1]; v[2] = C->edges[2 * bidx]; v[3] = C->edges[2 * bidx + 1];
          if (v[0] == v[2] || v[0] == v[3] || v[1] == v[2] || v[1] == v[3]) continue;
        }
        double X[12], P[12], cw[4], n[3];
        for (int j = 0; j < 4; ++j)
          for (int c = 0; c < 3; ++c) {
            X[3 * j + c] = (double)x[3 * v[j] + c];
            P[3 * j + c] = dxs ? (double)dxs[3 * v[j] + c] : 0.0;
          }
        if (mode == 2) {
          double t = orc_sc_accd(kind, X, P, cs);
          if (t < toi) toi = t;
          continue;
        }
        double d = orc_sc_distance(kind, X, cw, n);
        if (!(d < dh)) continue;
        if (!(d > 0)) { E += 1e300; continue; }
        double e, de, d2e;
        ipc64(d, dh, &e, &de, &d2e);
        E += kap * e;
        if (mode == 0) {
          double h = kap * fmax(d2e, 0.0);
          double nn[6] = {n[0] * n[0], n[1] * n[1], n[2] * n[2], n[0] * n[1], n[0] * n[2], n[1] * n[2]};
          for (int j = 0; j < 4; ++j) {
            for (int c = 0; c < 3; ++c) {
#pragma omp atomic
              C->sc_g[3 * v[j] + c] += kap * de * cw[j] * n[c];
            }
            for (int q = 0; q < 6; ++q) {
#pragma omp atomic
              C->sc_h[6 * v[j] + q] += h * cw[j] * cw[j] * nn[q];
            }
          }
        }
      }
  }
  free(bp); free(bt); free(be);
  return mode == 2 ? toi : E;
}

/* the pass functions for the tests (x, dxs: the oracle's real type) */
EXPORT double orc_sc_energy(orc_ctx* C, const real* x) { return sc_pass(C, x, NULL, 1); }
EXPORT double orc_sc_gradient(orc_ctx* C, const real* x) { return sc_pass(C, x, NULL, 0); }
EXPORT double orc_sc_toi(orc_ctx* C, const real* x, const real* dxs) { return sc_pass(C, x, dxs, 2); }

/* Shell.refresh (shell_energy.py:262-293): every triangle's W and rest
   area from the rest positions.  Call after the rest positions or the
   triangles change. */
EXPORT void orc_shell_refresh(orc_ctx* C) {
  for (int t = 0; t < C->ntris; ++t) {
    const int* tr = C->tris + 3 * t;
    C->tarea[t] = orc_tri_rest_frame(C->rest + 3 * tr[0], C->rest + 3 * tr[1], C->rest + 3 * tr[2], C->tW + 4 * t);
  }
}

/* shell energy per triangle into eS, gradient per vertex into gS, at x
   (Shell.evaluate, shell_energy.py:295-314); returns the summed energy */
static double shell_eval(orc_ctx* C, const real* x, int with_grad) {
  double sum = 0;
  real mu = C->p[P_SHELL_MU];
  if (with_grad) memset(C->gS, 0, sizeof(real) * 3 * (size_t)C->nv);
  for (int t = 0; t < C->ntris; ++t) {
    const int* tr = C->tris + 3 * t;
    real gr[9];
    real E = orc_membrane(x + 3 * tr[0], x + 3 * tr[1], x + 3 * tr[2], C->tW + 4 * t, C->tarea[t], mu,
                          with_grad ? gr : NULL);
    if (with_grad) {
      C->eS[t] = E;
      for (int a = 0; a < 3; ++a)
        for (int r = 0; r < 3; ++r) C->gS[3 * tr[a] + r] += gr[3 * a + r];
    }
    sum += (double)E;
  }
  return sum;
}

EXPORT void orc_update_system(orc_ctx* C, real dt) {
  int nv = C->nv, nt = C->nt;
  real idt2 = 1 / (dt * dt);
  const real* qb = C->p + P_QB0;
  real k = C->p[P_KAPPA], d0 = C->p[P_D0], d1 = C->p[P_D1];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nv; ++i) {
    for (int a = 0; a < 3; ++a) C->gx[3 * i + a] = C->mass[i] * (C->x[3 * i + a] - C->xt[3 * i + a]) * idt2;
    real d, gr[3], hs[9];
    int sid;
    vertex_distance(C, C->x + 3 * i, &d, gr, hs, &sid);
    C->dist[i] = d;
    C->sid[i] = sid;
    real b[3] = {0, 0, 0};
    if (d < d1) orc_barrier(d, d0, d1, qb, b);
    C->eB[i] = k * b[0];
    real h2 = b[2] > 0 ? b[2] : 0;
    for (int a = 0; a < 3; ++a) {
      C->gB[3 * i + a] = d < d1 ? k * b[1] * gr[a] : 0;
      for (int c = 0; c < 3; ++c) C->HB[9 * i + 3 * a + c] = d < d1 ? k * h2 * gr[a] * gr[c] : 0;
    }
    friction_at(C, i, C->x + 3 * i, dt, C->eD + i, C->gD + 3 * i, C->HD + 9 * i);
  }
  for (int pk = 0; pk < C->npins; ++pk) {
    int i = C->pin_idx[pk];
    for (int a = 0; a < 3; ++a) C->gx[3 * i + a] += C->p[P_ATTACH_K] * (C->x[3 * i + a] - C->pin_pos[3 * pk + a]) * idt2;
  }
  /* triangle vs capsule contact, folded into the vertex barrier arrays */
  if (C->p[P_TRI_CONTACT] != 0 && C->ntris > 0 && C->nshapes > 0) {
    for (int t = 0; t < C->ntris; ++t) {
      const int* tr = C->tris + 3 * t;
      real dd, w[3], n[3];
      int sid;
      tri_capsule(C, C->x + 3 * tr[0], C->x + 3 * tr[1], C->x + 3 * tr[2], &dd, w, n, &sid);
      C->tdist[t] = dd;
      memcpy(C->tbary + 3 * t, w, sizeof w);
      memcpy(C->tnorm + 3 * t, n, sizeof n);
      C->tsid[t] = sid;
      real b[3] = {0, 0, 0};
      if (dd < d1) orc_barrier(dd, d0, d1, qb, b);
      for (int a = 0; a < 3; ++a) {
        int vi = tr[a];
        C->eB[vi] += w[a] * k * b[0];
        for (int r = 0; r < 3; ++r) C->gB[3 * vi + r] += w[a] * k * b[1] * n[r];
      }
    }
  }
  if (C->p[P_SC] != 0) sc_pass(C, C->x, NULL, 0);
  if (C->p[P_SHELL_MU] > 0) shell_eval(C, C->x, 1);
  int energy = (int)C->p[P_ENERGY];
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nt; ++e) {
    const real* s = C->s + 6 * e;
    real mu = C->mat[3 * e], la = C->mat[3 * e + 1], vol = C->vol[e];
    if (energy == 0) orc_arap(s, mu, vol, C->psi + e, C->gs + 6 * e, C->Hs + 36 * e, C->Hsinv + 36 * e);
    else orc_neohookean(s, mu, la, vol, C->psi + e, C->gs + 6 * e, C->Hs + 36 * e, C->Hsinv + 36 * e, NULL);
    real xe[12];
    for (int a = 0; a < 4; ++a)
      for (int kk = 0; kk < 3; ++kk) xe[3 * a + kk] = C->x[3 * C->tets[4 * e + a] + kk];
    orc_tet_constraints(xe, C->dminv + 9 * e, s, C->c + 6 * e, C->G + 72 * e, NULL);
  }
}

/* _assemble_system (solver.py:641-799) */
EXPORT void orc_assemble(orc_ctx* C, real dt) {
  int nv = C->nv, nt = C->nt;
  real idt2 = 1 / (dt * dt);
  memset(C->val, 0, sizeof(real) * 9 * C->nnzb);
  for (int i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) C->g[3 * i + a] = -C->gx[3 * i + a];
  /* (built with OpenMP for the CPU baseline: tets in parallel, their
     contributions added atomically; the serial build sums in tet order) */
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nt; ++e) {
    real* Hl = C->Hl + 36 * e;
    const real* Hs = C->Hs + 36 * e;
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) Hl[a * 6 + b] = Hs[a * 6 + b] / (SYMD[a] * SYMD[b]);
    real t[6];
    for (int a = 0; a < 6; ++a) {
      real acc = 0;
      for (int b = 0; b < 6; ++b) acc += C->Hsinv[36 * e + a * 6 + b] * C->gs[6 * e + b];
      t[a] = acc;
    }
    for (int a = 0; a < 6; ++a) C->gl[6 * e + a] = C->c[6 * e + a] + SYMD[a] * t[a];
    const real* G = C->G + 72 * e;
    real HG[72], Ke[144], hg[6], fe[12];
    mm(Hl, G, HG, 6, 6, 12);
    for (int a = 0; a < 12; ++a)
      for (int b = 0; b < 12; ++b) {
        real acc = 0;
        for (int r = 0; r < 6; ++r) acc += G[r * 12 + a] * HG[r * 12 + b];
        Ke[a * 12 + b] = acc;
      }
    mm(Hl, C->gl + 6 * e, hg, 6, 6, 1);
    for (int a = 0; a < 12; ++a) {
      real acc = 0;
      for (int r = 0; r < 6; ++r) acc += G[r * 12 + a] * hg[r];
      fe[a] = acc;
    }
    for (int va = 0; va < 4; ++va) {
      int i = C->tets[4 * e + va];
      for (int r = 0; r < 3; ++r) {
#pragma omp atomic
        C->g[3 * i + r] += fe[3 * va + r];
      }
      for (int vb = 0; vb < 4; ++vb) {
        real* blk = C->val + 9 * C->tet_slot[16 * e + 4 * va + vb];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
#pragma omp atomic
            blk[3 * r + c] += Ke[(3 * va + r) * 12 + 3 * vb + c];
          }
      }
    }
  }
  for (int i = 0; i < nv; ++i) {
    real* blk = C->val + 9 * C->diag[i];
    real m = C->inv_mass[i] == 0 ? 1 : 1 / C->inv_mass[i];
    real kq = pin_of(C, i) >= 0 ? C->p[P_ATTACH_K] : 0;
    for (int r = 0; r < 3; ++r) {
      blk[4 * r] += (m + kq) * idt2;
      C->g[3 * i + r] -= C->gB[3 * i + r] + C->gD[3 * i + r];
    }
    for (int q = 0; q < 9; ++q) blk[q] += C->HB[9 * i + q] + C->HD[9 * i + q];
    if (C->p[P_SC] != 0) {
      /* self-contact gradient and diagonal block (sym6 order 00 11 22 01 02 12) */
      const double* h = C->sc_h + 6 * i;
      const int map[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
      for (int r = 0; r < 3; ++r) C->g[3 * i + r] -= (real)C->sc_g[3 * i + r];
      for (int q = 0; q < 9; ++q) blk[q] += (real)h[map[q]];
    }
  }
  /* triangle-capsule barrier Hessian blocks on the tri's vertex pairs; its
     energy and gradient were folded into eB / gB by orc_update_system
     (contact.py:684-775) */
  if (C->p[P_TRI_CONTACT] != 0 && C->ntris > 0 && C->nshapes > 0) {
    real k = C->p[P_KAPPA];
    for (int t = 0; t < C->ntris; ++t) {
      real d = C->tdist[t], b[3] = {0, 0, 0};
      if (d < C->p[P_D1]) orc_barrier(d, C->p[P_D0], C->p[P_D1], C->p + P_QB0, b);
      const real* w = C->tbary + 3 * t;
      const real* n = C->tnorm + 3 * t;
      real h2 = b[2] > 0 ? b[2] : 0;
      for (int a = 0; a < 3; ++a) {
        int vi = C->tris[3 * t + a];
        for (int bb = 0; bb < 3; ++bb) {
          int vj = C->tris[3 * t + bb];
          int lo = C->rowptr[vi], hi = C->rowptr[vi + 1] - 1, pos = -1;
          while (lo <= hi) {
            int mid = (lo + hi) / 2;
            if (C->col[mid] == vj) { pos = mid; break; }
            if (C->col[mid] < vj) lo = mid + 1; else hi = mid - 1;
          }
          if (pos < 0) continue;
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) C->val[9 * pos + 3 * r + c] += w[a] * w[bb] * k * h2 * n[r] * n[c];
        }
      }
    }
  }
  /* membrane shell: g -= grad, and the constant Gauss-Newton blocks
     mu A (w_a . w_b) I3, w_1 = row 0 of W, w_2 = row 1, w_0 = -(w_1 + w_2)
     (shell_energy.py:154-186, solver.py:789-797) */
  if (C->p[P_SHELL_MU] > 0) {
    real mu = C->p[P_SHELL_MU];
    for (int i = 0; i < nv; ++i)
      for (int r = 0; r < 3; ++r) C->g[3 * i + r] -= C->gS[3 * i + r];
    for (int t = 0; t < C->ntris; ++t) {
      const real* W = C->tW + 4 * t;
      real w[3][2] = {{-(W[0] + W[2]), -(W[1] + W[3])}, {W[0], W[1]}, {W[2], W[3]}};
      for (int a = 0; a < 3; ++a) {
        int vi = C->tris[3 * t + a];
        for (int bb = 0; bb < 3; ++bb) {
          int vj = C->tris[3 * t + bb];
          int lo = C->rowptr[vi], hi = C->rowptr[vi + 1] - 1, pos = -1;
          while (lo <= hi) {
            int mid = (lo + hi) / 2;
            if (C->col[mid] == vj) { pos = mid; break; }
            if (C->col[mid] < vj) lo = mid + 1; else hi = mid - 1;
          }
          if (pos < 0) continue;
          real h = mu * C->tarea[t] * (w[a][0] * w[bb][0] + w[a][1] * w[bb][1]);
          for (int r = 0; r < 3; ++r) C->val[9 * pos + 4 * r] += h;
        }
      }
    }
  }
}

EXPORT void orc_spmv(orc_ctx* C, const real* x, real* y) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < C->nv; ++i) {
    real acc[3] = {0, 0, 0};
    for (int k = C->rowptr[i]; k < C->rowptr[i + 1]; ++k) {
      const real* b = C->val + 9 * k;
      const real* xj = x + 3 * C->col[k];
      for (int r = 0; r < 3; ++r) acc[r] += b[3 * r] * xj[0] + b[3 * r + 1] * xj[1] + b[3 * r + 2] * xj[2];
    }
    y[3 * i] = acc[0]; y[3 * i + 1] = acc[1]; y[3 * i + 2] = acc[2];
  }
}

static double vdot(const real* a, const real* b, int n) {
  double acc = 0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
  for (int i = 0; i < n; ++i) acc += (double)a[i] * (double)b[i];
  return acc;
}

/* block-Jacobi inverse (kernels.py:32-53) + PCG on H dx = g, guess x_tilde - x.
   Returns CG iterations; rel-res (true, recomputed) in *relres. */
EXPORT int orc_global_solve(orc_ctx* C, double* relres) {
  int nv = C->nv, n = 3 * nv;
  for (int i = 0; i < nv; ++i) {
    real* M = C->Minv + 9 * i;
    const real* D = C->val + 9 * C->diag[i];
    if (C->p[P_PRECOND] == 0 || (real)fabs((double)det3(D)) < C->p[P_SING]) {
      for (int q = 0; q < 9; ++q) M[q] = (q % 4 == 0) ? 1 : 0;
    } else inv3(D, M);
  }
  /* initial guess: x~ - x (reference, solver.py:530-531); experiments: 1 = zero,
     2 = previous Newton step (kept in dx) */
  int gm = (int)C->p[P_GUESS];
  for (int k = 0; k < n; ++k) C->dx[k] = gm == 1 ? 0 : (gm == 2 ? C->dx[k] : C->xt[k] - C->x[k]);
  orc_spmv(C, C->dx, C->q);
  for (int k = 0; k < n; ++k) C->r[k] = C->g[k] - C->q[k];
  double bn = sqrt(vdot(C->g, C->g, n));
  double tol = (double)C->p[P_CG_TOL] * bn;
  int maxit = (int)C->p[P_CG_MAXIT], it = 0;
  for (int i = 0; i < nv; ++i) mm(C->Minv + 9 * i, C->r + 3 * i, C->z + 3 * i, 3, 3, 1);
  memcpy(C->pp, C->z, sizeof(real) * n);
  double rz = vdot(C->r, C->z, n);
  double rn = sqrt(vdot(C->r, C->r, n));
  while (rn > tol && it < maxit) {
    orc_spmv(C, C->pp, C->q);
    double pq = vdot(C->pp, C->q, n);
    if (pq == 0) break;
    real alpha = (real)(rz / pq);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k) { C->dx[k] += alpha * C->pp[k]; C->r[k] -= alpha * C->q[k]; }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nv; ++i) mm(C->Minv + 9 * i, C->r + 3 * i, C->z + 3 * i, 3, 3, 1);
    double rz2 = vdot(C->r, C->z, n);
    real beta = (real)(rz2 / rz);
    rz = rz2;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k) C->pp[k] = C->z[k] + beta * C->pp[k];
    rn = sqrt(vdot(C->r, C->r, n));
    ++it;
  }
  orc_spmv(C, C->dx, C->q);
  double acc = 0;
  for (int k = 0; k < n; ++k) { double d = (double)C->g[k] - (double)C->q[k]; acc += d * d; }
  if (relres) *relres = bn > 0 ? sqrt(acc) / bn : sqrt(acc);
  return it;
}

/* local solve (kernels.py:252-283); tet_lambda <- new lambda (solver.py:537) */
EXPORT void orc_local_solve(orc_ctx* C) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < C->nt; ++e) {
    real t[6];
    const real* G = C->G + 72 * e;
    for (int r = 0; r < 6; ++r) {
      real acc = C->gl[6 * e + r];
      for (int a = 0; a < 4; ++a) {
        const real* d = C->dx + 3 * C->tets[4 * e + a];
        acc -= G[r * 12 + 3 * a] * d[0] + G[r * 12 + 3 * a + 1] * d[1] + G[r * 12 + 3 * a + 2] * d[2];
      }
      t[r] = acc;
    }
    real lam[6];
    for (int r = 0; r < 6; ++r) {
      real acc = 0;
      for (int c = 0; c < 6; ++c) acc += C->Hl[36 * e + r * 6 + c] * t[c];
      lam[r] = -acc;
    }
    for (int r = 0; r < 6; ++r) {
      real acc = 0;
      for (int c = 0; c < 6; ++c) acc += C->Hsinv[36 * e + r * 6 + c] * (SYMD[c] * lam[c] + C->gs[6 * e + c]);
      C->ds[6 * e + r] = -acc;
    }
    memcpy(C->lam + 6 * e, lam, sizeof lam);
  }
}

/* merit f(x, s) with lambda fixed (solver.py:803-869) */
EXPORT double orc_merit(orc_ctx* C, const real* x, const real* s, real dt) {
  double idt2 = 1.0 / ((double)dt * (double)dt);
  double kin = 0, att = 0, ela = 0, bar = 0, fri = 0;
#pragma omp parallel for schedule(static) reduction(+ : kin)
  for (int i = 0; i < C->nv; ++i) {
    double d2 = 0;
    for (int a = 0; a < 3; ++a) { double d = (double)x[3 * i + a] - (double)C->xt[3 * i + a]; d2 += d * d; }
    kin += 0.5 * (double)C->mass[i] * d2;
  }
  for (int pk = 0; pk < C->npins; ++pk) {
    int i = C->pin_idx[pk];
    double d2 = 0;
    for (int a = 0; a < 3; ++a) { double d = (double)x[3 * i + a] - (double)C->pin_pos[3 * pk + a]; d2 += d * d; }
    att += 0.5 * (double)C->p[P_ATTACH_K] * d2;
  }
  int energy = (int)C->p[P_ENERGY];
#pragma omp parallel for schedule(static) reduction(+ : ela)
  for (int e = 0; e < C->nt; ++e) {
    const real* se = s + 6 * e;
    real mu = C->mat[3 * e], la = C->mat[3 * e + 1], vol = C->vol[e];
    double psi = energy == 0 ? (double)orc_arap_energy(se, mu, vol) : (double)orc_nh_energy(se, mu, la, vol);
    real xe[12], c[6];
    for (int a = 0; a < 4; ++a)
      for (int k = 0; k < 3; ++k) xe[3 * a + k] = x[3 * C->tets[4 * e + a] + k];
    orc_tet_constraints(xe, C->dminv + 9 * e, se, c, NULL, NULL);
    double lc = 0;
    for (int k = 0; k < 6; ++k) lc += (double)c[k] * (double)C->lam[6 * e + k];
    ela += psi - lc;
  }
  real kappa = C->p[P_KAPPA];
#pragma omp parallel for schedule(static) reduction(+ : bar, fri)
  for (int i = 0; i < C->nv; ++i) {
    real d;
    vertex_distance(C, x + 3 * i, &d, NULL, NULL, NULL);
    if (d < C->p[P_D1]) {
      real b[3];
      orc_barrier(d, C->p[P_D0], C->p[P_D1], C->p + P_QB0, b);
      bar += (double)(kappa * b[0]);
    }
    real E;
    friction_at(C, i, x + 3 * i, dt, &E, NULL, NULL);
    fri += E;
  }
  if (C->p[P_TRI_CONTACT] != 0 && C->ntris > 0 && C->nshapes > 0) {
#pragma omp parallel for schedule(static) reduction(+ : bar)
    for (int t = 0; t < C->ntris; ++t) {
      const int* tr = C->tris + 3 * t;
      real dd, w[3], n[3];
      int sid;
      tri_capsule(C, x + 3 * tr[0], x + 3 * tr[1], x + 3 * tr[2], &dd, w, n, &sid);
      if (dd < C->p[P_D1]) {
        real b[3];
        orc_barrier(dd, C->p[P_D0], C->p[P_D1], C->p + P_QB0, b);
        /* bary of the last full evaluation (contact.py:1050-1088) */
        const real* wl = C->tbary + 3 * t;
        bar += (double)(kappa * b[0]) * ((double)wl[0] + (double)wl[1] + (double)wl[2]);
      }
    }
  }
  if (C->p[P_SC] != 0) bar += sc_pass(C, x, NULL, 1);
  double shell = C->p[P_SHELL_MU] > 0 ? shell_eval(C, x, 0) : 0.0;
  return kin * idt2 + att * idt2 + ela + bar + fri + shell;
}

/* backtracking line search (line_search.py:76-163, solver.py:871-928).
   Returns the number of trials; *accepted is 1 when x, s were committed. */
EXPORT int orc_line_search(orc_ctx* C, real dt, int* accepted, double* alpha_out) {
  int nv = C->nv, nt = C->nt;
  /* slope t = -c (grad_x . dx + grad_s . ds) */
  double gdx = 0, gds = 0;
  real* gxf = C->q; /* scratch: g_x + gB + gD - G^T lambda */
  for (int k = 0; k < 3 * nv; ++k) gxf[k] = C->gx[k] + C->gB[k] + C->gD[k];
  if (C->p[P_SHELL_MU] > 0)
    for (int k = 0; k < 3 * nv; ++k) gxf[k] += C->gS[k];
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nt; ++e) {
    const real* G = C->G + 72 * e;
    for (int a = 0; a < 4; ++a) {
      int i = C->tets[4 * e + a];
      for (int r = 0; r < 3; ++r) {
        real acc = 0;
        for (int k = 0; k < 6; ++k) acc += G[k * 12 + 3 * a + r] * C->lam[6 * e + k];
#pragma omp atomic
        gxf[3 * i + r] -= acc;
      }
    }
  }
  gdx = vdot(gxf, C->dx, 3 * nv);
  if (C->p[P_SC] != 0)
    for (int k = 0; k < 3 * nv; ++k) gdx += C->sc_g[k] * (double)C->dx[k];
#pragma omp parallel for schedule(static) reduction(+ : gds)
  for (int e = 0; e < nt; ++e)
    for (int k = 0; k < 6; ++k) gds += ((double)C->gs[6 * e + k] + (double)SYMD[k] * (double)C->lam[6 * e + k]) * (double)C->ds[6 * e + k];
  double c = C->p[P_LS_C], tau = C->p[P_LS_TAU], thr = C->p[P_LS_THR];
  int maxit = (int)C->p[P_LS_MAX];
  double t = -c * gdx + -c * gds;
  double f0 = orc_merit(C, C->x, C->s, dt);
  double alpha = C->p[P_LS_A0];
  if (C->p[P_SC] != 0) {
    /* the first step cut to the self-contact time of impact */
    double toi = sc_pass(C, C->x, C->dx, 2);
    if (toi < alpha) alpha = (double)(float)toi;
  }
  int it = 0, cond;
  do {
    for (int k = 0; k < 3 * nv; ++k) C->xtrial[k] = C->x[k] + (real)alpha * C->dx[k];
    for (int k = 0; k < 6 * nt; ++k) C->strial[k] = C->s[k] + (real)alpha * C->ds[k];
    double f1 = orc_merit(C, C->xtrial, C->strial, dt);
    cond = ((f0 - f1) + thr < alpha * t) && it < maxit;
    ++it;
    if (alpha_out) *alpha_out = alpha;
    alpha *= tau;
  } while (cond);
  *accepted = it < maxit;
  if (*accepted) {
    memcpy(C->x, C->xtrial, sizeof(real) * 3 * nv);
    memcpy(C->s, C->strial, sizeof(real) * 6 * nt);
  }
  return it;
}

/* RefinementSolver.step without refinement (solver.py:429-579) */
EXPORT void orc_step(orc_ctx* C, real dt) {
  orc_begin_step(C, dt);
  int iters = (int)C->p[P_ITERS];
  C->stat_n = 0;
  for (int it = 0; it < iters; ++it) {
    orc_update_system(C, dt);
    orc_assemble(C, dt);
    double rr;
    int cgit = orc_global_solve(C, &rr);
    orc_local_solve(C);
    int acc = 1, lsit = 0;
    double alpha = 1;
    if (C->p[P_LS] != 0) lsit = orc_line_search(C, dt, &acc, &alpha);
    else {
      /* (with self-contact: the step cut to the time of impact) */
      real a = C->p[P_SC] != 0 ? (real)(float)sc_pass(C, C->x, C->dx, 2) : (real)1;
      for (int k = 0; k < 3 * C->nv; ++k) C->x[k] += a * C->dx[k];
      for (int k = 0; k < 6 * C->nt; ++k) C->s[k] += a * C->ds[k];
    }
    if (4 * it + 3 < 4096) {
      C->stats[4 * it + 0] = (real)cgit;
      C->stats[4 * it + 1] = (real)rr;
      C->stats[4 * it + 2] = (real)lsit;
      C->stats[4 * it + 3] = (real)(acc ? alpha : 0);
    }
    C->stat_n = it + 1;
  }
  for (int i = 0; i < 3 * C->nv; ++i) C->v[i] = (C->x[i] - C->xin[i]) / dt;
}

EXPORT void orc_destroy(orc_ctx* C) {
  if (!C) return;
  void* ptrs[] = {C->tets, C->dminv, C->mat, C->mass, C->inv_mass, C->vol, C->pin_idx, C->pin_pos, C->tris,
                  C->shape_type, C->shape_xf, C->shape_scale, C->shape_mu, C->rowptr, C->col, C->diag, C->tet_slot,
                  C->val, C->x, C->v, C->s, C->lam, C->xt, C->xstart, C->xin, C->lag_n, C->lag_fn, C->lag_sid,
                  C->c, C->G, C->gs, C->Hs, C->Hsinv, C->Hl, C->gl, C->psi, C->gx, C->gB, C->gD, C->HB, C->HD,
                  C->dist, C->ddist, C->hdist, C->sid, C->eB, C->eD, C->g, C->dx, C->ds, C->Minv, C->tdist,
                  C->tbary, C->tnorm, C->tsid, C->r, C->z, C->pp, C->q, C->xtrial, C->strial, C->pts, C->edges,
                  C->sc_g, C->sc_h, C->rest, C->tW, C->tarea, C->eS, C->gS};
  for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; ++k) free(ptrs[k]);
  free(C);
}

EXPORT int orc_real_size(void) { return (int)sizeof(real); }

/* Thread count of the OpenMP build (the serial build reports 1). */
EXPORT int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
