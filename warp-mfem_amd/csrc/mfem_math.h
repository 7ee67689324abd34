// Device math for the per-tet mixed-FEM element (gfx950, FP32 registers).
//
// Conventions (reference paths relative to the warp-mfem repository):
//   * 3x3 matrices are row-major float[9].
//   * sym6 order is (00, 11, 22, 01, 02, 12); SYM = diag(1,1,1,2,2,2)
//     (src/mfem/refinement/utils.py:4-20).
//   * F = Ds * DmInv with Ds = [x1-x0, x2-x0, x3-x0] as columns
//     (src/mfem/refinement/kernels.py:131-169).
//   * G (6x12) = -SYM * dS/dF * dF/dx, column 3k+i = vertex k coordinate i.
//
// Every loop has compile-time bounds and is unrolled so that all small
// arrays stay in VGPRs (no scratch).
#pragma once
#include <hip/hip_runtime.h>

#define MFEM_DEV __device__ __forceinline__

namespace mfem {

constexpr float SYMD[6] = {1.f, 1.f, 1.f, 2.f, 2.f, 2.f};

MFEM_DEV float det3(const float* A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}

MFEM_DEV void inv3(const float* A, float* B) {
  float id = 1.f / det3(A);
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

// ---------------------------------------------------------------------------
// 3x3 SVD, rotation variant.  Cyclic Jacobi on F^T F (4 fixed sweeps: the
// off-diagonal is below FP32 resolution after the fourth on deformation
// gradients from near-rest to 100 % strain, and a fifth changes nothing),
// sort descending, then QR of B = F V by Gram-Schmidt: U and V are rotations and
// the sign of det F lands on the smallest singular value, the convention of
// rotational_svd (src/mfem/refinement/utils.py:46-67).  Method: McAdams et
// al. 2011, "Computing the singular value decomposition of 3x3 matrices with
// minimal branching" (Jacobi + sort + QR), with plain Givens rotations.
// ---------------------------------------------------------------------------
// One Jacobi rotation of the symmetric A on (p, q), r the third index
// (Golub & Van Loan, Algorithm 8.4.1): tau = (a_qq - a_pp) / (2 a_pq),
// t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c;
// then a_pp -= t a_pq, a_qq += t a_pq, a_pq = 0, (a_rp, a_rq) rotated, and
// V's columns p, q rotated.  The reciprocals are the hardware's (1 ulp):
// an inexact t still gives an orthogonal rotation (c, s come from the same
// t), and what it leaves of a_pq is below FP32 rounding of the diagonal.
MFEM_DEV void jacobi_rot(float* A, float* V, int p, int q) {
  const int r = 3 - p - q;
  const float apq = A[p * 3 + q];
  // (the hardware reciprocal flushes denormals: 2 a_pq below the normal
  // range would give 0 * inf when a_pp == a_qq; such a rotation is a no-op
  // at FP32 precision anyway)
  if (!(fabsf(apq) > 1e-30f)) return;
  const float app = A[p * 3 + p], aqq = A[q * 3 + q];
  const float tau = (aqq - app) * __builtin_amdgcn_rcpf(2.f * apq);
  const float t = copysignf(1.f, tau) * __builtin_amdgcn_rcpf(fabsf(tau) + sqrtf(fmaf(tau, tau, 1.f)));
  const float cs = rsqrtf(fmaf(t, t, 1.f));
  const float sn = t * cs;
  A[p * 3 + p] = fmaf(-t, apq, app);
  A[q * 3 + q] = fmaf(t, apq, aqq);
  A[p * 3 + q] = A[q * 3 + p] = 0.f;
  const float arp = A[r * 3 + p], arq = A[r * 3 + q];
  const float nrp = cs * arp - sn * arq, nrq = sn * arp + cs * arq;
  A[r * 3 + p] = A[p * 3 + r] = nrp;
  A[r * 3 + q] = A[q * 3 + r] = nrq;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float vkp = V[k * 3 + p], vkq = V[k * 3 + q];
    V[k * 3 + p] = cs * vkp - sn * vkq;
    V[k * 3 + q] = sn * vkp + cs * vkq;
  }
}

MFEM_DEV void swap_cols(float* V, float* ev, int a, int b) {
  float t = ev[a];
  ev[a] = ev[b];
  ev[b] = t;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t = V[k * 3 + a];
    V[k * 3 + a] = V[k * 3 + b];
    V[k * 3 + b] = t;
  }
}

MFEM_DEV void svd3(const float* F, float* U, float* sig, float* V) {
  float A[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i * 3 + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
#pragma unroll
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.f : 0.f;
#pragma unroll
  for (int sweep = 0; sweep < 4; ++sweep) {
    jacobi_rot(A, V, 0, 1);
    jacobi_rot(A, V, 0, 2);
    jacobi_rot(A, V, 1, 2);
  }
  float ev[3] = {A[0], A[4], A[8]};
  if (ev[1] > ev[0]) swap_cols(V, ev, 0, 1);
  if (ev[2] > ev[0]) swap_cols(V, ev, 0, 2);
  if (ev[2] > ev[1]) swap_cols(V, ev, 1, 2);
  if (det3(V) < 0.f) {
    V[2] = -V[2];
    V[5] = -V[5];
    V[8] = -V[8];
  }
  float b0[3], b1[3], b2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    b0[k] = F[k * 3 + 0] * V[0] + F[k * 3 + 1] * V[3] + F[k * 3 + 2] * V[6];
    b1[k] = F[k * 3 + 0] * V[1] + F[k * 3 + 1] * V[4] + F[k * 3 + 2] * V[7];
    b2[k] = F[k * 3 + 0] * V[2] + F[k * 3 + 1] * V[5] + F[k * 3 + 2] * V[8];
  }
  float u0[3], u1[3], u2[3];
  float n0 = sqrtf(b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]);
  if (n0 > 1e-30f) {
    float in0 = 1.f / n0;
    u0[0] = b0[0] * in0; u0[1] = b0[1] * in0; u0[2] = b0[2] * in0;
  } else {
    u0[0] = 1.f; u0[1] = 0.f; u0[2] = 0.f;
  }
  float pr = u0[0] * b1[0] + u0[1] * b1[1] + u0[2] * b1[2];
  float w[3] = {b1[0] - pr * u0[0], b1[1] - pr * u0[1], b1[2] - pr * u0[2]};
  float n1 = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  if (n1 > 1e-6f * fmaxf(n0, 1e-30f)) {
    float in1 = 1.f / n1;
    u1[0] = w[0] * in1; u1[1] = w[1] * in1; u1[2] = w[2] * in1;
  } else {
    // orthogonal completion: cross u0 with the axis it is least aligned with
    float a0 = fabsf(u0[0]), a1 = fabsf(u0[1]), a2 = fabsf(u0[2]);
    float e[3] = {0.f, 0.f, 0.f};
    if (a0 <= a1 && a0 <= a2) e[0] = 1.f;
    else if (a1 <= a2) e[1] = 1.f;
    else e[2] = 1.f;
    u1[0] = u0[1] * e[2] - u0[2] * e[1];
    u1[1] = u0[2] * e[0] - u0[0] * e[2];
    u1[2] = u0[0] * e[1] - u0[1] * e[0];
    float in = rsqrtf(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    u1[0] *= in; u1[1] *= in; u1[2] *= in;
  }
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  sig[0] = u0[0] * b0[0] + u0[1] * b0[1] + u0[2] * b0[2];
  sig[1] = u1[0] * b1[0] + u1[1] * b1[1] + u1[2] * b1[2];
  sig[2] = u2[0] * b2[0] + u2[1] * b2[1] + u2[2] * b2[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    U[k * 3 + 0] = u0[k];
    U[k * 3 + 1] = u1[k];
    U[k * 3 + 2] = u2[k];
  }
}

// F = Ds * DmInv
MFEM_DEV void deformation_gradient(const float* x0, const float* x1, const float* x2, const float* x3,
                                   const float* D, float* F) {
  float e[3][3];  // e[k] = x_{k+1} - x0
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e[0][i] = x1[i] - x0[i];
    e[1][i] = x2[i] - x0[i];
    e[2][i] = x3[i] - x0[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) F[i * 3 + j] = e[0][i] * D[0 * 3 + j] + e[1][i] * D[1 * 3 + j] + e[2][i] * D[2 * 3 + j];
}

// S = V diag(sig) V^T as sym6
MFEM_DEV void stretch_sym6(const float* sig, const float* V, float* S6) {
  float S[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      S[i * 3 + j] = V[i * 3 + 0] * sig[0] * V[j * 3 + 0] + V[i * 3 + 1] * sig[1] * V[j * 3 + 1] +
                     V[i * 3 + 2] * sig[2] * V[j * 3 + 2];
  S6[0] = S[0]; S6[1] = S[4]; S6[2] = S[8];
  S6[3] = S[1]; S6[4] = S[2]; S6[5] = S[5];
}

// dS/dF (6x9, column 3i+j <-> F_ij) from the twist-mode form of dR/dF
// (src/mfem/refinement/utils.py:75-121):
//   dR/dF_ij = sum_k (2/s_k) t_k t_k[ij],  t_k = U T_k V^T / sqrt(2)
//   dS_ij    = F^T dR/dF_ij + e_j (row i of R)
MFEM_DEV void stretch_gradient(const float* F, const float* U, const float* sig, const float* V, float* dS) {
  // twist tensors M_k = U T_k V^T / sqrt(2): T_0 twists (0,1), T_1 (1,2), T_2 (0,2)
  float M[3][9];
  const float isq = 0.70710678118654752f;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      // U T V^T with T = e_a e_b^T - e_b e_a^T  ->  u_a v_b^T - u_b v_a^T (signs per T_k)
      // T_0 = [[0,-1,0],[1,0,0],[0,0,0]]: -u0 v1^T + u1 v0^T
      M[0][r * 3 + c] = isq * (U[r * 3 + 1] * V[c * 3 + 0] - U[r * 3 + 0] * V[c * 3 + 1]);
      // T_1 = [[0,0,0],[0,0,1],[0,-1,0]]: u1 v2^T - u2 v1^T
      M[1][r * 3 + c] = isq * (U[r * 3 + 1] * V[c * 3 + 2] - U[r * 3 + 2] * V[c * 3 + 1]);
      // T_2 = [[0,0,1],[0,0,0],[-1,0,0]]: u0 v2^T - u2 v0^T
      M[2][r * 3 + c] = isq * (U[r * 3 + 0] * V[c * 3 + 2] - U[r * 3 + 2] * V[c * 3 + 0]);
    }
  float co[3] = {2.f / fmaxf(sig[0] + sig[1], 1e-8f), 2.f / fmaxf(sig[1] + sig[2], 1e-8f),
                 2.f / fmaxf(sig[0] + sig[2], 1e-8f)};
  // FtM[k] = F^T M_k
  float FtM[3][9];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        FtM[k][r * 3 + c] = F[0 * 3 + r] * M[k][0 * 3 + c] + F[1 * 3 + r] * M[k][1 * 3 + c] + F[2 * 3 + r] * M[k][2 * 3 + c];
  float R[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) R[i * 3 + j] = U[i * 3 + 0] * V[j * 3 + 0] + U[i * 3 + 1] * V[j * 3 + 1] + U[i * 3 + 2] * V[j * 3 + 2];
  // sym6 entries (row, col) pairs
  constexpr int SR[6] = {0, 1, 2, 0, 0, 1};
  constexpr int SC[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float w0 = co[0] * M[0][i * 3 + j], w1 = co[1] * M[1][i * 3 + j], w2 = co[2] * M[2][i * 3 + j];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        int r = SR[q], c = SC[q];
        float v = w0 * FtM[0][r * 3 + c] + w1 * FtM[1][r * 3 + c] + w2 * FtM[2][r * 3 + c];
        if (r == j) v += R[i * 3 + c];
        dS[q * 9 + 3 * i + j] = v;
      }
    }
}

// G = -SYM dS/dF dF/dx (6x12 row-major); dF_ij/dx_{k+1,i} = D[k][j],
// dF_ij/dx_{0,i} = -sum_k D[k][j]  (src/mfem/refinement/utils.py:123-177)
MFEM_DEV void constraint_jacobian(const float* dS, const float* D, float* G) {
  float Dp[4][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    Dp[1][j] = D[0 * 3 + j];
    Dp[2][j] = D[1 * 3 + j];
    Dp[3][j] = D[2 * 3 + j];
    Dp[0][j] = -(D[0 * 3 + j] + D[1 * 3 + j] + D[2 * 3 + j]);
  }
#pragma unroll
  for (int q = 0; q < 6; ++q)
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        float v = dS[q * 9 + 3 * i + 0] * Dp[k][0] + dS[q * 9 + 3 * i + 1] * Dp[k][1] + dS[q * 9 + 3 * i + 2] * Dp[k][2];
        G[q * 12 + 3 * k + i] = -SYMD[q] * v;
      }
}

// ---------------------------------------------------------------------------
// energies in stretch space (src/mfem/refinement/elastic_energy.py)
// ---------------------------------------------------------------------------
MFEM_DEV float frob2_s(const float* s) {
  return s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2.f * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
}
MFEM_DEV float det_s(const float* s) {
  return (s[0] * s[1] * s[2] + 2.f * s[3] * s[4] * s[5]) - (s[1] * s[4] * s[4] + s[0] * s[5] * s[5] + s[2] * s[3] * s[3]);
}
MFEM_DEV void grad_det_s(const float* s, float* g) {
  g[0] = s[1] * s[2] - s[5] * s[5];
  g[1] = s[0] * s[2] - s[4] * s[4];
  g[2] = s[0] * s[1] - s[3] * s[3];
  g[3] = 2.f * (s[4] * s[5] - s[2] * s[3]);
  g[4] = 2.f * (s[3] * s[5] - s[1] * s[4]);
  g[5] = 2.f * (s[3] * s[4] - s[0] * s[5]);
}

// ARAP energy with the reference's I2 (elastic_energy.py:110-112, 147-149)
MFEM_DEV float arap_energy(const float* s, float mu, float vol) {
  float I1 = s[0] + s[1] + s[2];
  float I2 = s[0] * s[1] + s[1] * s[2] + s[2] * s[0] + 2.f * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  return vol * (mu * (I2 + 3.f - 2.f * I1) * 0.5f);
}
MFEM_DEV float nh_energy(const float* s, float mu, float lam, float vol) {
  float I2 = frob2_s(s), J = det_s(s);
  return vol * (mu * (I2 - 3.f) * 0.5f - mu * (J - 1.f) + lam * 0.5f * (J - 1.f) * (J - 1.f));
}

// GMW81 modified Cholesky (Nocedal & Wright Alg. 3.4 variant,
// src/mfem/numerical/gmw81.py:18-86) then A+E = L L^T and (A+E)^-1 = Z^T Z,
// Z = L^-1 (src/mfem/numerical/psd_fix.py:157-191).  n = 6, in registers.
MFEM_DEV void psd_fix6(float* A, float* Hfix, float* Hinv) {
  constexpr int n = 6;
  const float eps = 1.19e-7f, delta = 3.4496376e-4f;  // sqrt(1.19e-7)
  float gamma = 0.f, xi = 0.f;
#pragma unroll
  for (int i = 0; i < n; ++i) {
    gamma = fmaxf(gamma, fabsf(A[i * n + i]));
#pragma unroll
    for (int j = i + 1; j < n; ++j) xi = fmaxf(xi, fabsf(A[i * n + j]));
  }
  float b2 = fmaxf(gamma, fmaxf(xi / 5.9160798f /* sqrt(35) */, eps));
#pragma unroll
  for (int j = 0; j < n; ++j) {
#pragma unroll
    for (int i = j + 1; i < n; ++i) {
      float cij = A[j * n + i];
#pragma unroll
      for (int k = 0; k < j; ++k) cij -= A[k * n + k] * A[k * n + k] * A[i * n + k] * A[j * n + k];
      A[i * n + j] = cij;
    }
    float theta = 0.f;
#pragma unroll
    for (int i = j + 1; i < n; ++i) theta = fmaxf(theta, fabsf(A[i * n + j]));
    float dj = fmaxf(delta, fmaxf(fabsf(A[j * n + j]), theta * theta / b2));
    A[j * n + j] = sqrtf(dj);
    float idj = 1.f / dj;
#pragma unroll
    for (int i = j + 1; i < n; ++i) {
      float cij = A[i * n + j];
      A[i * n + j] = cij * idj;
      A[i * n + i] -= cij * cij * idj;
    }
  }
  float L[36];
#pragma unroll
  for (int j = 0; j < n; ++j) {
    float sd = A[j * n + j];
#pragma unroll
    for (int i = 0; i < n; ++i) L[i * n + j] = (i > j) ? A[i * n + j] * sd : (i == j ? sd : 0.f);
  }
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k <= j; ++k) acc += L[i * n + k] * L[j * n + k];
      Hfix[i * n + j] = acc;
      Hfix[j * n + i] = acc;
    }
#pragma unroll
  for (int k = 0; k < n; ++k) {
    L[k * n + k] = 1.f / L[k * n + k];
#pragma unroll
    for (int i = k + 1; i < n; ++i) {
      float acc = 0.f;
#pragma unroll
      for (int j = k; j < i; ++j) acc += L[i * n + j] * L[j * n + k];
      L[i * n + k] = -acc / L[i * n + i];
    }
  }
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      float acc = 0.f;
#pragma unroll
      for (int k = i; k < n; ++k) acc += L[k * n + i] * L[k * n + j];
      Hinv[i * n + j] = acc;
      Hinv[j * n + i] = acc;
    }
}

// Neo-Hookean gradient and PSD-fixed Hessian (elastic_energy.py:255-346)
MFEM_DEV void nh_grad_hess(const float* s, float mu, float lam, float vol, float* g, float* H, float* Hinv) {
  float J = det_s(s), gJ[6];
  grad_det_s(s, gJ);
  float a = lam * (J - 1.f) - mu;
#pragma unroll
  for (int k = 0; k < 6; ++k) g[k] = vol * (mu * SYMD[k] * s[k] + a * gJ[k]);
  float hJ[36] = {0.f, s[2], s[1], 0.f, 0.f, -2.f * s[5],
                  s[2], 0.f, s[0], 0.f, -2.f * s[4], 0.f,
                  s[1], s[0], 0.f, -2.f * s[3], 0.f, 0.f,
                  0.f, 0.f, -2.f * s[3], -2.f * s[2], 2.f * s[5], 2.f * s[4],
                  0.f, -2.f * s[4], 0.f, 2.f * s[5], -2.f * s[1], 2.f * s[3],
                  -2.f * s[5], 0.f, 0.f, 2.f * s[4], 2.f * s[3], -2.f * s[0]};
  float h[36];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) h[i * 6 + j] = vol * ((i == j ? mu * SYMD[i] : 0.f) + lam * gJ[i] * gJ[j] + a * hJ[i * 6 + j]);
  psd_fix6(h, H, Hinv);
}

MFEM_DEV void arap_grad_hess(const float* s, float mu, float vol, float* g, float* H, float* Hinv) {
#pragma unroll
  for (int k = 0; k < 6; ++k) g[k] = vol * mu * (SYMD[k] * s[k] - (k < 3 ? 1.f : 0.f));
#pragma unroll
  for (int i = 0; i < 36; ++i) {
    H[i] = 0.f;
    Hinv[i] = 0.f;
  }
  float imv = 1.f / (mu * vol);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    H[k * 7] = mu * vol * SYMD[k];
    Hinv[k * 7] = imv * (k < 3 ? 1.f : 0.5f);
  }
}

// ---------------------------------------------------------------------------
// contact (src/mfem/refinement/contact.py:16-59, src/mfem/ipc/distance.py)
// ---------------------------------------------------------------------------
MFEM_DEV void barrier(float d, float d0, float d1, const float* q, float& e, float& de, float& d2e) {
  e = 0.f; de = 0.f; d2e = 0.f;
  if (d < d1 && d > d0) {
    float l = logf(d / d1), dd = d - d1;
    e = -dd * dd * l;
    de = -(2.f * dd * l + dd * dd / d);
    d2e = -(2.f * l + 4.f * dd / d - dd * dd / (d * d));
  } else if (d <= d0) {
    float t = d - d0;
    e = q[0] + q[1] * t + 0.5f * q[2] * t * t;
    de = q[1] + q[2] * t;
    d2e = q[2];
  }
}
MFEM_DEV float friction_f0(float y, float ev) {
  if (y >= ev) return y;
  return -y * y * y / (3.f * ev * ev) + y * y / ev + ev / 3.f;
}
MFEM_DEV float friction_f1_over_y(float y, float ev) {
  if (y >= ev) return 1.f / fmaxf(y, 1e-12f);
  return -y / (ev * ev) + 2.f / ev;
}
template <typename T, typename Q>
MFEM_DEV void quat_to_mat(const Q* q, T* R) {
  const T x = q[0], y = q[1], z = q[2], w = q[3], one = 1, two = 2;
  R[0] = one - two * (y * y + z * z); R[1] = two * (x * y - z * w); R[2] = two * (x * z + y * w);
  R[3] = two * (x * y + z * w); R[4] = one - two * (x * x + z * z); R[5] = two * (y * z - x * w);
  R[6] = two * (x * z - y * w); R[7] = two * (y * z + x * w); R[8] = one - two * (x * x + y * y);
}

// barrier(d) in FP64 (distances inside the barrier range): the log branch on
// (d0, d1) and, for d <= d0, its quadratic extension from d0 with the
// coefficients formed here in FP64
MFEM_DEV void barrier64(double d, double d0, double d1, double& e, double& de, double& d2e) {
  e = 0.0; de = 0.0; d2e = 0.0;
  const double x = d > d0 ? d : d0;
  // (no extension without a positive d0, as the host's coefficients)
  if (!(x < d1) || !(x > 0.0)) return;
  const double l = log(x / d1), dd = x - d1;
  e = -dd * dd * l;
  de = -(2.0 * dd * l + dd * dd / x);
  d2e = -(2.0 * l + 4.0 * dd / x - dd * dd / (x * x));
  if (d <= d0) {
    const double t = d - d0;
    e += de * t + 0.5 * d2e * t * t;
    de += d2e * t;
  }
}

// ---------------------------------------------------------------------------
// triangle / segment closest points (triangle-vs-capsule contact,
// src/mfem/refinement/contact.py:417-550)
// ---------------------------------------------------------------------------
MFEM_DEV float dot3f(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
MFEM_DEV void cross3f(const float* a, const float* b, float* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
MFEM_DEV T dot3t(const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <typename T>
MFEM_DEV void cross3t(const T* a, const T* b, T* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
template <typename T>
MFEM_DEV T tmin(T a, T b) { return a < b ? a : b; }
template <typename T>
MFEM_DEV T tmax(T a, T b) { return a > b ? a : b; }

// (templated on the scalar type: FP32 everywhere, FP64 for contact distances
// inside the barrier range)
// barycentric weights of the point of triangle abc closest to x (Voronoi
// regions of the vertices, then the edges, then the face; Ericson,
// Real-Time Collision Detection 5.1.5)
template <typename T>
MFEM_DEV void closest_point_triangle_bary(const T* x, const T* a, const T* b, const T* c, T* w) {
  T ab[3], ac[3], ax[3], bx[3], cx[3];
  for (int i = 0; i < 3; ++i) {
    ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ax[i] = x[i] - a[i]; bx[i] = x[i] - b[i]; cx[i] = x[i] - c[i];
  }
  const T d1 = dot3t(ab, ax), d2 = dot3t(ac, ax);
  if (d1 <= T(0) && d2 <= T(0)) { w[0] = T(1); w[1] = T(0); w[2] = T(0); return; }
  const T d3 = dot3t(ab, bx), d4 = dot3t(ac, bx);
  if (d3 >= T(0) && d4 <= d3) { w[0] = T(0); w[1] = T(1); w[2] = T(0); return; }
  const T d5 = dot3t(ab, cx), d6 = dot3t(ac, cx);
  if (d6 >= T(0) && d5 <= d6) { w[0] = T(0); w[1] = T(0); w[2] = T(1); return; }
  const T vc = d1 * d4 - d3 * d2;
  if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
    const T v = d1 / (d1 - d3);
    w[0] = T(1) - v; w[1] = v; w[2] = T(0);
    return;
  }
  const T vb = d5 * d2 - d1 * d6;
  if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
    const T t = d2 / (d2 - d6);
    w[0] = T(1) - t; w[1] = T(0); w[2] = t;
    return;
  }
  const T va = d3 * d6 - d5 * d4;
  if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
    const T t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    w[0] = T(0); w[1] = T(1) - t; w[2] = t;
    return;
  }
  const T tot = va + vb + vc;
  if (tot <= T(0)) { w[0] = w[1] = w[2] = T(1) / T(3); return; }
  const T v = vb / tot, t = vc / tot;
  w[0] = T(1) - v - t; w[1] = v; w[2] = t;
}

// closest points of segments p1q1 and p2q2 (Ericson 5.1.9); returns the
// parameter on the first segment (t_out, if given: on the second)
template <typename T>
MFEM_DEV T segment_segment(const T* p1, const T* q1, const T* p2, const T* q2, T* c1, T* c2,
                           T* t_out = nullptr) {
  const T d1[3] = {q1[0] - p1[0], q1[1] - p1[1], q1[2] - p1[2]};
  const T d2[3] = {q2[0] - p2[0], q2[1] - p2[1], q2[2] - p2[2]};
  const T r[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const T a = dot3t(d1, d1), e = dot3t(d2, d2), f = dot3t(d2, r);
  const T eps = T(1e-12);
  T s, t;
  if (a <= eps && e <= eps) {
    s = T(0); t = T(0);
  } else if (a <= eps) {
    s = T(0); t = tmin(tmax(f / e, T(0)), T(1));
  } else {
    const T c = dot3t(d1, r);
    if (e <= eps) {
      t = T(0); s = tmin(tmax(-c / a, T(0)), T(1));
    } else {
      const T b = dot3t(d1, d2), den = a * e - b * b;
      s = den != T(0) ? tmin(tmax((b * f - c * e) / den, T(0)), T(1)) : T(0);
      t = (b * s + f) / e;
      if (t < T(0)) {
        t = T(0); s = tmin(tmax(-c / a, T(0)), T(1));
      } else if (t > T(1)) {
        t = T(1); s = tmin(tmax((b - c) / a, T(0)), T(1));
      }
    }
  }
  for (int i = 0; i < 3; ++i) {
    c1[i] = p1[i] + d1[i] * s;
    c2[i] = p2[i] + d2[i] * t;
  }
  if (t_out) *t_out = t;
  return s;
}

// does segment p0p1 pierce triangle abc (Moller-Trumbore with the parameter
// limited to [0, 1])?  w = barycentric weights of the piercing point
template <typename T>
MFEM_DEV bool segment_triangle_intersect(const T* p0, const T* p1, const T* a, const T* b,
                                         const T* c, T* w) {
  T dir[3], e1[3], e2[3], pv[3], tv[3], qv[3];
  for (int i = 0; i < 3; ++i) {
    dir[i] = p1[i] - p0[i]; e1[i] = b[i] - a[i]; e2[i] = c[i] - a[i]; tv[i] = p0[i] - a[i];
  }
  cross3t(dir, e2, pv);
  const T det = dot3t(e1, pv);
  const T scale = sqrt(dot3t(dir, dir)) * sqrt(dot3t(e1, e1)) * sqrt(dot3t(e2, e2));
  if (fabs(det) <= T(1e-7) * scale + T(1e-30)) return false;
  const T inv = T(1) / det;
  const T u = dot3t(tv, pv) * inv;
  if (u < T(0) || u > T(1)) return false;
  cross3t(tv, e1, qv);
  const T v = dot3t(dir, qv) * inv;
  if (v < T(0) || u + v > T(1)) return false;
  const T t = dot3t(e2, qv) * inv;
  if (t < T(0) || t > T(1)) return false;
  w[0] = T(1) - u - v; w[1] = u; w[2] = v;
  return true;
}

// distance from triangle abc to segment p0p1: the two endpoints against the
// face, the three edges against the segment, then the pierce test (distance
// 0).  bary = weights of the triangle point, q = the segment point.
template <typename T>
MFEM_DEV T closest_point_triangle_segment(const T* a, const T* b, const T* c, const T* p0,
                                              const T* p1, T* bary, T* q, bool* hit) {
  T w[3], p[3];
  closest_point_triangle_bary(p0, a, b, c, bary);
  for (int i = 0; i < 3; ++i) {
    p[i] = bary[0] * a[i] + bary[1] * b[i] + bary[2] * c[i] - p0[i];
    q[i] = p0[i];
  }
  T best = sqrt(dot3t(p, p));
  closest_point_triangle_bary(p1, a, b, c, w);
  for (int i = 0; i < 3; ++i) p[i] = w[0] * a[i] + w[1] * b[i] + w[2] * c[i] - p1[i];
  T d = sqrt(dot3t(p, p));
  if (d < best) {
    best = d;
    for (int i = 0; i < 3; ++i) { bary[i] = w[i]; q[i] = p1[i]; }
  }
  const T* E0[3] = {a, b, c};
  const T* E1[3] = {b, c, a};
  for (int e = 0; e < 3; ++e) {
    T c1[3], c2[3];
    const T s = segment_segment(E0[e], E1[e], p0, p1, c1, c2);
    const T df[3] = {c1[0] - c2[0], c1[1] - c2[1], c1[2] - c2[2]};
    d = sqrt(dot3t(df, df));
    if (d < best) {
      best = d;
      if (e == 0) { bary[0] = T(1) - s; bary[1] = s; bary[2] = T(0); }
      else if (e == 1) { bary[0] = T(0); bary[1] = T(1) - s; bary[2] = s; }
      else { bary[0] = s; bary[1] = T(0); bary[2] = T(1) - s; }
      for (int i = 0; i < 3; ++i) q[i] = c2[i];
    }
  }
  *hit = segment_triangle_intersect(p0, p1, a, b, c, w);
  if (*hit) {
    best = T(0);
    for (int i = 0; i < 3; ++i) bary[i] = w[i];
    for (int i = 0; i < 3; ++i) q[i] = w[0] * a[i] + w[1] * b[i] + w[2] * c[i];
  }
  return best;
}

}  // namespace mfem
