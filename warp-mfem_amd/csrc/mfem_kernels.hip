// mfem_kernels.hip -- gfx950 kernels + C ABI for the per-timestep soft-body
// solve (warp-mfem RefinementSolver.step, src/mfem/refinement/solver.py:429-579).
//
// Kernel map (one launch each unless noted; all use 256-thread workgroups):
//   k_begin_step   per vertex   predictor, friction anchor, lagged normal force
//   k_element      per tet      F, SVD, c, G, energy, PSD Hessian, Schur blocks,
//                               12x12 stiffness (10 unique 3x3 blocks) and force;
//                  per vertex   (its last workgroups) kinetic/attachment
//                               gradient, barrier, friction, shell gradient
//   k_assemble     per lane-slot of the SELL-64 matrix: deterministic gather of
//                               the tet blocks that land on that block; the
//                               right-hand side (gather of tet forces) and the
//                               diagonal mass/contact terms of each row
//   k_pcg          persistent, cooperative: whole block-Jacobi PCG loop with
//                               FP64 reductions and register-resident row state
//   k_local        per tet      lambda, ds (stage path; with fused_local the
//                               line search's first pass does it)
//   k_linesearch   persistent, cooperative: whole backtracking loop, FP64 merit
//   k_end_step     per vertex   velocity
//   (self-contact, mfem_selfcontact.hip: candidates, pair barrier, CCD)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "../../include/mfem_amd.h"
#include "mfem_math.h"
#include "mfem_selfcontact.h"

using namespace mfem;

#define MFEM_VERSION 1
static constexpr int kBlock = 256;
static constexpr int kWaves = kBlock / 64;
// Grid-barrier wait limit in wall_clock64 ticks (100 MHz on CDNA3/4): a wait
// this long means a workgroup never arrived; the kernel then records the error
// and exits instead of spinning forever.
static constexpr uint64_t kBarrierTimeout = 500000000ull;
// A grid-barrier wait looks at the error bits (and the limit above) only
// after this many ticks (20 us), so that a normal wait costs one memory
// round trip per poll.
static constexpr uint64_t kStatusCheckAfter = 2000ull;
// partials read per lane per batch of poll loads (6 x 32 = 192 workgroups)
static constexpr int kPollPasses = 6;

static thread_local std::string g_last_error;
static thread_local int g_last_pcg_variant = -1;  // the variant the last mfem_global_solve launched

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                                   \
  do {                                                                                                  \
    hipError_t _e = (expr);                                                                             \
    if (_e != hipSuccess) return fail(MFEM_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e));   \
  } while (0)

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load3(const float* p, int i, float* o) {
  o[0] = p[3 * i];
  o[1] = p[3 * i + 1];
  o[2] = p[3 * i + 2];
}
// packed upper triangle index of (i, j), i <= j, for a 6x6 symmetric matrix
__device__ __forceinline__ constexpr int sym21(int i, int j) {
  return i <= j ? (i * 6 - i * (i - 1) / 2 + (j - i)) : (j * 6 - j * (j - 1) / 2 + (i - j));
}
// the 10 unique vertex pairs of a tet: 4 diagonal then the 6 upper pairs
__device__ constexpr int PA[10] = {0, 1, 2, 3, 0, 0, 0, 1, 1, 2};
__device__ constexpr int PB[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};

// Distances inside the barrier range are formed in FP64 (positions and poses
// are FP32 inputs): in FP32 the subtraction of ~1 m coordinates leaves an
// absolute error of ~1e-7 m, a relative error of ~1e-3 at d = 1e-4 m.  Shapes
// are ranked in FP32; the winner's distance is recomputed in FP64 when the
// FP32 value is below kFp64Range d1.
static constexpr float kFp64Range = 2.0f;

// FP64 signed distance of x to shape k (plane: local z; capsule: distance to
// the axis segment minus the radius) (src/mfem/ipc/distance.py:4-57)
__device__ __forceinline__ double shape_distance64(const mfem_solver& S, int k, const float* x) {
  const float* xf = S.shape_xf + 7 * k;
  double R[9];
  quat_to_mat(xf + 3, R);
  const double dw[3] = {(double)x[0] - (double)xf[0], (double)x[1] - (double)xf[1], (double)x[2] - (double)xf[2]};
  double xl[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) xl[i] = R[i] * dw[0] + R[3 + i] * dw[1] + R[6 + i] * dw[2];
  if (S.shape_type[k] == MFEM_SHAPE_PLANE) return xl[2];
  const double rad = S.shape_scale[3 * k], hh = S.shape_scale[3 * k + 1];
  const double az = fmin(fmax(xl[2], -hh), hh);
  const double df[3] = {xl[0], xl[1], xl[2] - az};
  return sqrt(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]) - rad;
}

// minimum signed distance over the rigid shapes, world gradient, shape id
// (src/mfem/refinement/contact.py:62-164, src/mfem/ipc/distance.py:4-57);
// d64 (may be null) receives the FP64 distance of the nearest shape (the FP32
// value when it is outside kFp64Range d1)
__device__ __forceinline__ float shape_distance(const mfem_solver& S, const float* x, float* grad, int* sid,
                                double* d64 = nullptr) {
  float dm = 1.0e8f;
  grad[0] = grad[1] = grad[2] = 0.f;
  *sid = -1;
  for (int k = 0; k < S.n_shapes; ++k) {
    const float* xf = S.shape_xf + 7 * k;
    float R[9];
    quat_to_mat(xf + 3, R);
    float dw[3] = {x[0] - xf[0], x[1] - xf[1], x[2] - xf[2]};
    float xl[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) xl[i] = R[i] * dw[0] + R[3 + i] * dw[1] + R[6 + i] * dw[2];
    float d, n[3] = {0.f, 0.f, 1.f};
    if (S.shape_type[k] == MFEM_SHAPE_PLANE) {
      d = xl[2];
    } else if (S.shape_type[k] == MFEM_SHAPE_CAPSULE) {
      float rad = S.shape_scale[3 * k], hh = S.shape_scale[3 * k + 1];
      float az = fminf(fmaxf(xl[2], -hh), hh);
      float df[3] = {xl[0], xl[1], xl[2] - az};
      float rho = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
      if (rho > 1e-8f) {
        n[0] = df[0] / rho; n[1] = df[1] / rho; n[2] = df[2] / rho;
      }
      d = rho - rad;
    } else {
      continue;
    }
    if (d < dm) {
      dm = d;
#pragma unroll
      for (int i = 0; i < 3; ++i) grad[i] = R[i * 3] * n[0] + R[i * 3 + 1] * n[1] + R[i * 3 + 2] * n[2];
      *sid = k;
    }
  }
  if (d64) *d64 = (*sid >= 0 && dm < kFp64Range * S.d1) ? shape_distance64(S, *sid, x) : (double)dm;
  return dm;
}

// lagged IPC friction at position x (contact.py:304-387)
__device__ __forceinline__ float friction_eval(const mfem_solver& S, int i, const float* x, float dt, float* grad, float* hess) {
  float lamn = S.lag_fn[i];
  int sid = S.lag_sid[i];
  if (grad) grad[0] = grad[1] = grad[2] = 0.f;
  if (hess)
#pragma unroll
    for (int q = 0; q < 9; ++q) hess[q] = 0.f;
  if (!(lamn > 0.f && sid >= 0)) return 0.f;
  float mu = S.shape_mu[sid];
  if (!(mu > 0.f)) return 0.f;
  float n[3], xs[3];
  load3(S.lag_n, i, n);
  load3(S.x_start, i, xs);
  float u[3] = {x[0] - xs[0], x[1] - xs[1], x[2] - xs[2]};
  float un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
  float ut[3] = {u[0] - un * n[0], u[1] - un * n[1], u[2] - un * n[2]};
  float y = sqrtf(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
  float ev = S.eps_v * dt, sc = mu * lamn;
  float f1y = friction_f1_over_y(y, ev);
  if (grad)
#pragma unroll
    for (int k = 0; k < 3; ++k) grad[k] = sc * f1y * ut[k];
  if (hess)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) hess[a * 3 + b] = sc * f1y * ((a == b ? 1.f : 0.f) - n[a] * n[b]);
  return sc * friction_f0(y, ev);
}

// distance of triangle abc to capsule k minus its radius, the barycentric
// weights of the closest triangle point and the contact normal (capsule ->
// triangle; the face normal oriented away from the axis midpoint when the
// axis touches or pierces the face) (contact.py:553-681), in precision T
template <typename T>
__device__ __forceinline__ T tri_capsule_one(const mfem_solver& S, int k, const T* a, const T* b, const T* c, T* bary, T* nrm) {
  const float* xf = S.shape_xf + 7 * k;
  T R[9];
  quat_to_mat(xf + 3, R);
  const T rad = S.shape_scale[3 * k], hh = S.shape_scale[3 * k + 1];
  T p0[3], p1[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    p0[i] = (T)xf[i] - R[i * 3 + 2] * hh;
    p1[i] = (T)xf[i] + R[i * 3 + 2] * hh;
  }
  T q[3];
  bool hit;
  const T dist = closest_point_triangle_segment(a, b, c, p0, p1, bary, q, &hit);
  T p[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) p[i] = bary[0] * a[i] + bary[1] * b[i] + bary[2] * c[i];
  if (!hit && dist > T(1e-8)) {
#pragma unroll
    for (int i = 0; i < 3; ++i) nrm[i] = (p[i] - q[i]) / dist;
  } else {
    const T e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const T e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    cross3t(e1, e2, nrm);
    const T nl = sqrt(dot3t(nrm, nrm));
    if (nl > T(1e-20)) {
      nrm[0] /= nl; nrm[1] /= nl; nrm[2] /= nl;
    } else {
      nrm[0] = 0; nrm[1] = 0; nrm[2] = 1;
    }
    const T ax[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const T pp0[3] = {p[0] - p0[0], p[1] - p0[1], p[2] - p0[2]};
    const T aa = dot3t(ax, ax);
    const T t = dot3t(pp0, ax) / (aa > T(1e-20) ? aa : T(1e-20));
    const T sgn = t < T(0.5) ? T(-1) : T(1);
    if (sgn * dot3t(nrm, ax) < T(0)) {
      nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; nrm[2] = -nrm[2];
    }
  }
  return dist - rad;
}

// nearest capsule of triangle abc: ranked in FP32, the winner recomputed in
// FP64 inside kFp64Range d1 (weights, normal, distance); d64 (may be null) =
// that distance in FP64
__device__ __forceinline__ float tri_capsule(const mfem_solver& S, const float* a, const float* b, const float* c, float* bary,
                             float* nrm, int* sid, double* d64 = nullptr) {
  float dmin = 1.0e8f;
  bary[0] = bary[1] = bary[2] = 1.f / 3.f;
  nrm[0] = nrm[1] = nrm[2] = 0.f;
  *sid = -1;
  for (int k = 0; k < S.n_shapes; ++k) {
    if (S.shape_type[k] != MFEM_SHAPE_CAPSULE) continue;
    float w[3], n[3];
    const float d = tri_capsule_one<float>(S, k, a, b, c, w, n);
    if (d < dmin) {
      dmin = d;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        bary[i] = w[i];
        nrm[i] = n[i];
      }
      *sid = k;
    }
  }
  double dd = dmin;
  if (*sid >= 0 && dmin < kFp64Range * S.d1) {
    double A[3], B[3], C[3], w[3], n[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      A[i] = a[i]; B[i] = b[i]; C[i] = c[i];
    }
    dd = tri_capsule_one<double>(S, *sid, A, B, C, w, n);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      bary[i] = (float)w[i];
      nrm[i] = (float)n[i];
    }
    dmin = (float)dd;
  }
  if (d64) *d64 = dd;
  return dmin;
}

__device__ __forceinline__ bool tri_contact_on(const mfem_solver& S) {
  return S.tri_contact && S.n_tris > 0 && S.n_shapes > 0;
}

// ---------------------------------------------------------------------------
// membrane shell (src/mfem/refinement/shell_energy.py)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool shell_on(const mfem_solver& S) { return S.shell_on && S.n_tris > 0; }

// columns f1, f2 of the 3x2 F = [x1 - x0, x2 - x0] W of triangle t at
// x + a dx (shell_energy.py:66-80)
__device__ __forceinline__ void shell_F(const mfem_solver& S, int t, float a, const float* W, float* f1, float* f2) {
  const int i0 = S.tris[3 * t], i1 = S.tris[3 * t + 1], i2 = S.tris[3 * t + 2];
  float x0[3], d1[3], d2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x0[k] = S.x[3 * i0 + k] + a * (float)S.dx[3 * i0 + k];
    d1[k] = S.x[3 * i1 + k] + a * (float)S.dx[3 * i1 + k] - x0[k];
    d2[k] = S.x[3 * i2 + k] + a * (float)S.dx[3 * i2 + k] - x0[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f1[k] = d1[k] * W[0] + d2[k] * W[2];
    f2[k] = d1[k] * W[1] + d2[k] * W[3];
  }
}

// mu A |F - R|^2 / 2 = mu A (|F|^2 - 2 (sigma1 + sigma2) + 2) / 2 with
// sigma1 + sigma2 = sqrt(tr C + 2 sqrt(det C)), C = F^T F
// (shell_energy.py:83-96); FP64 from the FP32 columns (the merit sums it
// against terms of similar size)
__device__ __forceinline__ double shell_energy64(const float* f1, const float* f2, float mu, float area) {
  const double c11 = (double)f1[0] * f1[0] + (double)f1[1] * f1[1] + (double)f1[2] * f1[2];
  const double c12 = (double)f1[0] * f2[0] + (double)f1[1] * f2[1] + (double)f1[2] * f2[2];
  const double c22 = (double)f2[0] * f2[0] + (double)f2[1] * f2[1] + (double)f2[2] * f2[2];
  const double dc = fmax(c11 * c22 - c12 * c12, 0.0);
  const double ssum = sqrt(fmax(c11 + c22 + 2.0 * sqrt(dc), 1e-20));
  return (double)mu * (double)area * 0.5 * (c11 + c22 - 2.0 * ssum + 2.0);
}

// gradient of corner `corner` of a triangle: mu A (F - R) W^T mapped to the
// corner, with F - R = F (I - S^-1), S = (C + sqrt(det C) I) / (sigma1 +
// sigma2), zero stress when sqrt(det C) < 1e-10 (shell_energy.py:99-125,
// 190-229)
__device__ __forceinline__ void shell_corner_grad(const float* f1, const float* f2, const float* W, float mu,
                                                  float area, int corner, float* g) {
  const float c11 = f1[0] * f1[0] + f1[1] * f1[1] + f1[2] * f1[2];
  const float c12 = f1[0] * f2[0] + f1[1] * f2[1] + f1[2] * f2[2];
  const float c22 = f2[0] * f2[0] + f2[1] * f2[1] + f2[2] * f2[2];
  const float sd = sqrtf(fmaxf(c11 * c22 - c12 * c12, 0.f));
  const float den = sqrtf(fmaxf(c11 + c22 + 2.f * sd, 1e-20f));
  if (sd < 1e-10f) {
    g[0] = g[1] = g[2] = 0.f;
    return;
  }
  const float sc = mu * area, inv = 1.f / (den * sd);
  const float i11 = (c22 + sd) * inv, i12 = -c12 * inv, i22 = (c11 + sd) * inv;
  // corner 1 takes row 0 of W, corner 2 row 1, corner 0 minus their sum
  const float wa = corner == 1 ? W[0] : corner == 2 ? W[2] : -(W[0] + W[2]);
  const float wb = corner == 1 ? W[1] : corner == 2 ? W[3] : -(W[1] + W[3]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float p1 = sc * (f1[k] - (f1[k] * i11 + f2[k] * i12));
    const float p2 = sc * (f2[k] - (f1[k] * i12 + f2[k] * i22));
    g[k] = p1 * wa + p2 * wb;
  }
}

// per triangle: W = Dm^-1 in the in-plane frame u = e1 / |e1|, v = n x u,
// and the rest area; both 0 for a degenerate rest triangle
// (shell_energy.py:37-63, 137-152)
__global__ void __launch_bounds__(kBlock) k_shell_frames(mfem_solver S) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= S.n_tris) return;
  float r0[3], r1[3], r2[3];
  load3(S.rest, S.tris[3 * t], r0);
  load3(S.rest, S.tris[3 * t + 1], r1);
  load3(S.rest, S.tris[3 * t + 2], r2);
  const float e1[3] = {r1[0] - r0[0], r1[1] - r0[1], r1[2] - r0[2]};
  const float e2[3] = {r2[0] - r0[0], r2[1] - r0[1], r2[2] - r0[2]};
  float n[3];
  cross3f(e1, e2, n);
  const float ta = sqrtf(dot3f(n, n));
  float W[4] = {0.f, 0.f, 0.f, 0.f}, area = 0.f;
  if (!(ta < 1e-12f)) {
    const float le = sqrtf(dot3f(e1, e1));
    const float u[3] = {e1[0] / le, e1[1] / le, e1[2] / le};
    const float nn[3] = {n[0] / ta, n[1] / ta, n[2] / ta};
    float v[3];
    cross3f(nn, u, v);
    const float a = dot3f(e1, u), b = dot3f(e2, u), c = dot3f(e1, v), d = dot3f(e2, v);
    const float det = a * d - b * c;
    W[0] = d / det;
    W[1] = -b / det;
    W[2] = -c / det;
    W[3] = a / det;
    area = 0.5f * ta;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) S.tri_W[4 * t + k] = W[k];
  S.tri_area[t] = area;
}

// per lane-slot: the constant Gauss-Newton coefficient sum over the
// triangle vertex pairs (a, b) on this block of mu A (w_a . w_b), w_1 = row 0
// of W, w_2 = row 1, w_0 = -(w_1 + w_2) (shell_energy.py:154-186)
__global__ void __launch_bounds__(kBlock) k_shell_hess(mfem_solver S) {
  const int l = blockIdx.x * kBlock + threadIdx.x;
  if (l >= S.n_slots * 64) return;
  float acc = 0.f;
  for (int k = S.tcon_ptr[l]; k < S.tcon_ptr[l + 1]; ++k) {
    const int src = S.tcon_src[k], t = src / 9, a = (src / 3) % 3, b = src % 3;
    const float* W = S.tri_W + 4 * t;
    const float w[3][2] = {{-(W[0] + W[2]), -(W[1] + W[3])}, {W[0], W[1]}, {W[2], W[3]}};
    acc += S.shell_mu * S.tri_area[t] * (w[a][0] * w[b][0] + w[a][1] * w[b][1]);
  }
  S.shell_h[l] = acc;
}

// shell energy of triangle t at x + a dx (shell_energy.py:216-229)
__device__ __forceinline__ double shell_merit(const mfem_solver& S, int t, float a) {
  const float* W = S.tri_W + 4 * t;
  const float Wl[4] = {W[0], W[1], W[2], W[3]};
  float f1[3], f2[3];
  shell_F(S, t, a, Wl, f1, f2);
  return shell_energy64(f1, f2, S.shell_mu, S.tri_area[t]);
}

// ---------------------------------------------------------------------------
// per-vertex kernels
// ---------------------------------------------------------------------------

// per surface triangle: nearest capsule, closest point, normal
__global__ void __launch_bounds__(kBlock) k_tri_contact(mfem_solver S) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= S.n_tris) return;
  const int i0 = S.tris[3 * t], i1 = S.tris[3 * t + 1], i2 = S.tris[3 * t + 2];
  float a[3], b[3], c[3], w[3], n[3];
  load3(S.x, i0, a);
  load3(S.x, i1, b);
  load3(S.x, i2, c);
  int sid;
  const float d = tri_capsule(S, a, b, c, w, n, &sid);
  S.tri_dist[t] = d;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    S.tri_bary[3 * t + k] = w[k];
    S.tri_nrm[3 * t + k] = n[k];
  }
  S.tri_sid[t] = sid;
}
__global__ void __launch_bounds__(kBlock) k_begin_step(mfem_solver S, float dt) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.n_verts) return;
  float x[3], v[3];
  load3(S.x, i, x);
  load3(S.v, i, v);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    S.x_start[3 * i + k] = x[k];
    S.x_tilde[3 * i + k] = x[k] + v[k] * dt + S.gravity[k] * dt * dt / 2.0f;
  }
  if (!S.friction_on) return;
  float grad[3];
  int sid;
  double d;
  shape_distance(S, x, grad, &sid, &d);
  float lamn = 0.f, n[3] = {0.f, 0.f, 0.f};
  int s = -1;
  if (d < S.d1) {
    double e, de, d2e;
    barrier64(d, S.d0, S.d1, e, de, d2e);
    lamn = (float)((double)S.kappa * fmax(-de, 0.0));
    n[0] = grad[0]; n[1] = grad[1]; n[2] = grad[2];
    s = sid;
  }
  S.lag_n[3 * i] = n[0];
  S.lag_n[3 * i + 1] = n[1];
  S.lag_n[3 * i + 2] = n[2];
  S.lag_fn[i] = lamn;
  S.lag_sid[i] = s;
}

// per vertex: kinetic + attachment gradient, vertex and triangle barrier,
// friction, shell gradient (run by the element kernel's vertex workgroups)
__device__ __forceinline__ void vertex_update(const mfem_solver& S, int i, float dt) {
  const float idt2 = 1.f / (dt * dt);
  float x[3], xt[3];
  load3(S.x, i, x);
  load3(S.x_tilde, i, xt);
  float m = S.mass[i];
  float gx[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) gx[k] = m * (x[k] - xt[k]) * idt2;
  if (S.inv_mass[i] == 0.f) {
    float p[3];
    load3(S.pin_pos, i, p);
#pragma unroll
    for (int k = 0; k < 3; ++k) gx[k] += S.attach_k * (x[k] - p[k]) * idt2;
  }
  float grad[3] = {0.f, 0.f, 0.f}, eb = 0.f, gb[3] = {0.f, 0.f, 0.f}, H[9];
  int sid;
  double d = 1.0e8;
  if (S.n_shapes > 0) shape_distance(S, x, grad, &sid, &d);
  float h2 = 0.f;
  if (d < S.d1) {
    double e, de, d2e;
    barrier64(d, S.d0, S.d1, e, de, d2e);
    const double kap = S.kappa;
    eb = (float)(kap * e);
#pragma unroll
    for (int k = 0; k < 3; ++k) gb[k] = (float)(kap * de) * grad[k];
    h2 = (float)(kap * fmax(d2e, 0.0));
  }
  if (tri_contact_on(S)) {
    // triangle-capsule barrier energy and gradient of the incident
    // triangles, in incidence order (contact.py:684-775)
    for (int kk = S.vtri_ptr[i]; kk < S.vtri_ptr[i + 1]; ++kk) {
      const int src = S.vtri_src[kk], t = src / 3, a = src - 3 * t;
      const float dt_ = S.tri_dist[t];
      if (!(dt_ < S.d1)) continue;
      double e, de, d2e;
      barrier64(dt_, S.d0, S.d1, e, de, d2e);
      const float wa = S.tri_bary[3 * t + a];
      eb += wa * (float)((double)S.kappa * e);
      const float kde = (float)((double)S.kappa * de);
#pragma unroll
      for (int k = 0; k < 3; ++k) gb[k] += wa * kde * S.tri_nrm[3 * t + k];
    }
  }
  if (shell_on(S)) {
    // membrane shell gradient of the incident triangles, in incidence order
    float gsh[3] = {0.f, 0.f, 0.f};
    for (int kk = S.vtri_ptr[i]; kk < S.vtri_ptr[i + 1]; ++kk) {
      const int src = S.vtri_src[kk], t = src / 3, a = src - 3 * t;
      const float W[4] = {S.tri_W[4 * t], S.tri_W[4 * t + 1], S.tri_W[4 * t + 2], S.tri_W[4 * t + 3]};
      float f1[3], f2[3], gc[3];
      shell_F(S, t, 0.f, W, f1, f2);
      shell_corner_grad(f1, f2, W, S.shell_mu, S.tri_area[t], a, gc);
#pragma unroll
      for (int k = 0; k < 3; ++k) gsh[k] += gc[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) S.gS[3 * i + k] = gsh[k];
  }
  float gd[3], hd[9];
  float ed = friction_eval(S, i, x, dt, gd, hd);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) H[a * 3 + b] = h2 * grad[a] * grad[b] + hd[a * 3 + b];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    S.gx[3 * i + k] = gx[k];
    S.gB[3 * i + k] = gb[k];
    S.gD[3 * i + k] = gd[k];
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) S.Hv[9 * i + q] = H[q];
  S.dist[i] = (float)d;
  S.eB[i] = eb;
  S.eD[i] = ed;
}

// ---------------------------------------------------------------------------
// per-tet element kernel
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gather_tet(const mfem_solver& S, int e, float (&xv)[4][3], int4& tv) {
  tv = reinterpret_cast<const int4*>(S.tets)[e];
  load3(S.x, tv.x, xv[0]);
  load3(S.x, tv.y, xv[1]);
  load3(S.x, tv.z, xv[2]);
  load3(S.x, tv.w, xv[3]);
}

// Wave-cooperative store of whole per-tet records of N floats (record e at
// arr + N e) from registers, through this wave's LDS slab: a record-per-lane
// store touches one 64-byte line per lane and instruction, the wave's records
// are one contiguous span, so lane t stores elements t, t + 64, ... of the
// span (a few lines per instruction) after the slab has transposed them.  Slab rows have an odd stride (no bank conflicts).  LDS
// operations of a wave complete in order.  nvalid = records of the wave that
// exist (the others are not written).
constexpr int slab_stride(int n) { return n | 1; }
constexpr int kSlabFloats = 64 * 73;  // rows of up to 72 floats
template <int N>
__device__ __forceinline__ void wave_store(float* slab, float* dst, const float* v, int nvalid, int lane) {
  constexpr int W = slab_stride(N);
  static_assert(64 * W <= kSlabFloats, "slab size");
#pragma unroll
  for (int q = 0; q < N; ++q) slab[lane * W + q] = v[q];
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int t = lane + 64 * k;
    const int l = t / N;
    if (l < nvalid) dst[t] = slab[l * W + (t - l * N)];
  }
  __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(kBlock) k_init_stretch(mfem_solver S) {
  int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= S.n_tets) return;
  float xv[4][3];
  int4 tv;
  gather_tet(S, e, xv, tv);
  float D[9], F[9], U[9], sg[3], V[9], S6[6];
#pragma unroll
  for (int q = 0; q < 9; ++q) D[q] = S.dminv[9 * e + q];
  deformation_gradient(xv[0], xv[1], xv[2], xv[3], D, F);
  svd3(F, U, sg, V);
  stretch_sym6(sg, V, S6);
#pragma unroll
  for (int k = 0; k < 6; ++k) S.s[6 * e + k] = S6[k];
}

// Per-tet element quantities at the current positions (kernels.py:131-249,
// elastic_energy.py): constraint c = SYM (sym6(S) - s), its Jacobian
// G = -SYM dS/dF dF/dx (6x12), sym6(S), the elastic energy, gradient and
// PSD-fixed Hessian and its inverse in stretch space, and the Schur blocks
// H_l = INV_SYM H INV_SYM, g_l = c + SYM H^-1 g_s.
struct TetElement {
  float G[72], c[6], S6[6], gs[6], Hi[36], Hl[36], gl[6], psi;
};

__device__ __forceinline__ void tet_element(const mfem_solver& S, int e, const float (&xv)[4][3], const float* D,
                                            const float* s, TetElement& o) {
  const float mu = S.materials[3 * e], la = S.materials[3 * e + 1], vol = S.vol[e];
  float F[9], U[9], sg[3], V[9], dS[54];
  deformation_gradient(xv[0], xv[1], xv[2], xv[3], D, F);
  svd3(F, U, sg, V);
  stretch_sym6(sg, V, o.S6);
#pragma unroll
  for (int k = 0; k < 6; ++k) o.c[k] = SYMD[k] * (o.S6[k] - s[k]);
  stretch_gradient(F, U, sg, V, dS);
  constraint_jacobian(dS, D, o.G);
  float H[36];
  if ((S.tet_energy ? S.tet_energy[e] : S.energy) == 0) {
    o.psi = arap_energy(s, mu, vol);
    arap_grad_hess(s, mu, vol, o.gs, H, o.Hi);
  } else {
    o.psi = nh_energy(s, mu, la, vol);
    nh_grad_hess(s, mu, la, vol, o.gs, H, o.Hi);
  }
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 6; ++b) o.Hl[a * 6 + b] = H[a * 6 + b] / (SYMD[a] * SYMD[b]);
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    float acc = 0.f;
#pragma unroll
    for (int b = 0; b < 6; ++b) acc += o.Hi[a * 6 + b] * o.gs[b];
    o.gl[a] = o.c[a] + SYMD[a] * acc;
  }
}

// Element kernel.  FULL: also stores G, H_l, g_l, H_s^-1, g_s for the stage
// local solve (k_local) and the line search's slope; otherwise only what
// assembly needs (the 10 unique blocks of K = G^T H_l G, the force
// G^T H_l g_l) and the energy, the local solve then running inside the line
// search (k_linesearch, fused_local).
template <bool FULL>
__global__ void __launch_bounds__(kBlock, 2) k_element(mfem_solver S, float dt, int n_tet_blocks) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  if ((int)blockIdx.x >= n_tet_blocks) {
    // the vertex workgroups (independent of the tets' work)
    const int i = ((int)blockIdx.x - n_tet_blocks) * kBlock + threadIdx.x;
    if (i < S.n_verts) vertex_update(S, i, dt);
    return;
  }
  // per-tet outputs leave through wave_store; lanes past the last tet compute
  // the last tet again and store nothing
  constexpr int kSlab = FULL ? kSlabFloats : 64 * 13;
  __shared__ float slabs[kWaves][kSlab];
  const int lane = threadIdx.x & 63;
  const int e0 = blockIdx.x * kBlock + (threadIdx.x & ~63);
  const int nvalid = min(64, S.n_tets - e0);
  if (nvalid <= 0) return;  // (the whole wave)
  const int e = min(e0 + lane, S.n_tets - 1);
  float* slab = slabs[threadIdx.x >> 6];
  float xv[4][3];
  int4 tv;
  gather_tet(S, e, xv, tv);
  float D[9], s[6];
#pragma unroll
  for (int q = 0; q < 9; ++q) D[q] = S.dminv[9 * e + q];
#pragma unroll
  for (int q = 0; q < 6; ++q) s[q] = S.s[6 * e + q];
  TetElement el;
  tet_element(S, e, xv, D, s, el);
  float hg[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    float acc = 0.f;
#pragma unroll
    for (int b = 0; b < 6; ++b) acc += el.Hl[a * 6 + b] * el.gl[b];
    hg[a] = acc;
  }
  if constexpr (FULL) {
    // stores needed by the stage local solve and the line search
    wave_store<72>(slab, S.G + (size_t)72 * e0, el.G, nvalid, lane);
    float h21[21], i21[21];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = i; j < 6; ++j) {
        h21[sym21(i, j)] = el.Hl[i * 6 + j];
        i21[sym21(i, j)] = el.Hi[i * 6 + j];
      }
    wave_store<21>(slab, S.Hl + (size_t)21 * e0, h21, nvalid, lane);
    wave_store<21>(slab, S.Hsinv + (size_t)21 * e0, i21, nvalid, lane);
    wave_store<6>(slab, S.gl + (size_t)6 * e0, el.gl, nvalid, lane);
    wave_store<6>(slab, S.gs + (size_t)6 * e0, el.gs, nvalid, lane);
  }
  if (lane < nvalid) S.psi[e] = el.psi;
  // HG = H_l G  (6x12)
  float HG[72];
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      float acc = 0.f;
#pragma unroll
      for (int b = 0; b < 6; ++b) acc += el.Hl[a * 6 + b] * el.G[b * 12 + k];
      HG[a * 12 + k] = acc;
    }
  // force f = G^T H_l g_l and the 10 unique blocks of K = G^T H_l G
  {
    float fe[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      float acc = 0.f;
#pragma unroll
      for (int a = 0; a < 6; ++a) acc += el.G[a * 12 + k] * hg[a];
      fe[k] = acc;
    }
    wave_store<12>(slab, S.fe + (size_t)12 * e0, fe, nvalid, lane);
  }
  const size_t T = (size_t)S.n_tets;
#pragma unroll
  for (int p = 0; p < 10; ++p) {
    float kb[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        float acc = 0.f;
#pragma unroll
        for (int a = 0; a < 6; ++a) acc += el.G[a * 12 + 3 * PA[p] + r] * HG[a * 12 + 3 * PB[p] + cc];
        kb[3 * r + cc] = acc;
      }
    wave_store<9>(slab, S.Kb + ((size_t)p * T + e0) * 9, kb, nvalid, lane);
  }
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
__device__ __forceinline__ size_t val_index(int lane_slot, int q) {
  return ((size_t)(lane_slot >> 6) * 9 + q) * 64 + (lane_slot & 63);
}

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round-robin
// by workgroup index.  This bijection of [0, gridDim.x) gives each XCD one
// contiguous range of blocks instead of every eighth block, so that gathers
// from neighbouring rows (records of shared tets) meet in the same L2.
__device__ __forceinline__ int xcd_block() {
  constexpr int kXcds = 8;
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds, k = b / kXcds;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Right-hand side of vertex i (solver.py:737-797): g = sum of its tets'
// forces G^T H_l g_l (in incidence order) - g_x - grad B - grad D
// [- self-contact - shell]
__device__ __forceinline__ void vertex_g(const mfem_solver& S, int i) {
  float g[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) g[k] = -S.gx[3 * i + k];
  const int k0 = S.vinc_ptr[i], k1 = S.vinc_ptr[i + 1];
  for (int k = k0; k < k1; ++k) {
    int src = S.vinc_src[k];
    const float* f = S.fe + (size_t)(src >> 2) * 12 + 3 * (src & 3);
    g[0] += f[0];
    g[1] += f[1];
    g[2] += f[2];
  }
  float gsc[3] = {0.f, 0.f, 0.f};
  if (self_contact_on(S))
#pragma unroll
    for (int k = 0; k < 3; ++k) gsc[k] = (float)S.sc_g[3 * i + k];
  if (shell_on(S))
#pragma unroll
    for (int k = 0; k < 3; ++k) gsc[k] += S.gS[3 * i + k];
#pragma unroll
  for (int k = 0; k < 3; ++k) S.g[3 * i + k] = g[k] - S.gB[3 * i + k] - S.gD[3 * i + k] - gsc[k];
}

// additions to vertex i's diagonal block: (M + Q) / dt^2 + H_B + H_D
// [+ the self-contact block] (solver.py:737-797)
__device__ __forceinline__ void vertex_diag(const mfem_solver& S, int i, float dt, float (&diag)[9]) {
  float hsc[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (self_contact_on(S)) {
    const double* h = S.sc_h + 6 * (size_t)i;
    const float h6[6] = {(float)h[0], (float)h[1], (float)h[2], (float)h[3], (float)h[4], (float)h[5]};
    hsc[0] = h6[0]; hsc[4] = h6[1]; hsc[8] = h6[2];
    hsc[1] = hsc[3] = h6[3];
    hsc[2] = hsc[6] = h6[4];
    hsc[5] = hsc[7] = h6[5];
  }
  const float idt2 = 1.f / (dt * dt);
  float im = S.inv_mass[i];
  float md = (im == 0.f ? 1.f + S.attach_k : 1.f / im) * idt2;
#pragma unroll
  for (int q = 0; q < 9; ++q) diag[q] = S.Hv[9 * i + q] + ((q % 4 == 0) ? md : 0.f) + hsc[q];
}

// Assembly of H into the SELL-64 pattern, one thread per lane-slot, and of
// the right-hand side: the lanes of each chunk's first slot form g of their
// rows, and each row's diagonal lane-slot adds its vertex terms before the
// block is stored.
__global__ void __launch_bounds__(kBlock) k_assemble(mfem_solver S, float dt) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  int l = xcd_block() * kBlock + threadIdx.x;
  if (l >= S.n_slots * 64) return;
  float acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.f;
  const int k0 = S.contrib_ptr[l], k1 = S.contrib_ptr[l + 1];
  for (int k = k0; k < k1; ++k) {
    int src = S.contrib_src[k];
    const float* B = S.Kb + (size_t)(src >> 1) * 9;
    float b[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) b[q] = B[q];
    if (src & 1) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[3 * r + c] += b[3 * c + r];
    } else {
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] += b[q];
    }
  }
  if (tri_contact_on(S)) {
    // w_a w_b kappa max(b'', 0) n n^T of each triangle on this block
    for (int k = S.tcon_ptr[l]; k < S.tcon_ptr[l + 1]; ++k) {
      const int src = S.tcon_src[k], t = src / 9, a = (src / 3) % 3, bb = src % 3;
      const float d = S.tri_dist[t];
      if (!(d < S.d1)) continue;
      double e, de, d2e;
      barrier64(d, S.d0, S.d1, e, de, d2e);
      const float h = S.tri_bary[3 * t + a] * S.tri_bary[3 * t + bb] * (float)((double)S.kappa * fmax(d2e, 0.0));
      const float* n = S.tri_nrm + 3 * t;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[3 * r + c] += h * n[r] * n[c];
    }
  }
  if (shell_on(S)) {
    // the membrane shell's constant Gauss-Newton block (value I3)
    const float h = S.shell_h[l];
    acc[0] += h;
    acc[4] += h;
    acc[8] += h;
  }
  // the row of this lane-slot: chunk c holds slots [chunk_ptr[c], chunk_ptr[c + 1])
  const int slot = l >> 6;
  int lo = 0, hi = S.n_chunks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S.chunk_ptr[mid] <= slot) lo = mid;
    else hi = mid - 1;
  }
  const int row = lo * 64 + (l & 63);
  if (row < S.n_verts) {
    if (slot == S.chunk_ptr[lo]) vertex_g(S, row);
    if (S.diag_lane[row] == l) {
      float add[9];
      vertex_diag(S, row, dt, add);
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] += add[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) S.sell_val[val_index(l, q)] = acc[q];
}

// ---------------------------------------------------------------------------
// grid-wide reduction + barrier for the persistent kernels
//
// Cross-workgroup data (published vectors, partial sums) is written and read
// with agent-scope coherent accesses (sc1: write-through to memory on store,
// no stale L1/L2 hit on load).  Arrival protocol: every wave waits for its
// coherent stores to complete (s_waitcnt vmcnt(0)), the workgroup
// synchronises, and 2 * NV threads store the workgroup's FP64 partials as
// epoch-tagged 64-bit halves.  Every workgroup polls all tagged partials
// (arrival and value in one load), then sums them in one fixed order, so
// every workgroup obtains the bit-identical total and takes the same branch.
// This avoids agent-scope release/acquire fences, whose L2 write-back /
// invalidate cost ~13 us per barrier at ~180 workgroups on MI355X.  Partials
// ping-pong between two slots by epoch parity.  Waits are bounded
// (kBarrierTimeout): on timeout the error bit is set and every workgroup
// leaves the kernel.
// ---------------------------------------------------------------------------
struct GridSync {
  double* red;
  uint32_t* flags;
  int32_t* status;
  uint32_t epoch;
  bool abort;
  int wg0, wg1;   // the workgroups that take part (one body's)
  size_t region;  // word offset of this kernel's partials in red
};

// largest b in [0, n) with ptr[b] <= x (ptr ascending, ptr[0] == 0)
__device__ __forceinline__ int body_of(const int32_t* ptr, int n, int x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ptr[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
constexpr size_t kRedRegionWords = (size_t)2 * MFEM_RED_MAX_BLOCKS * 8;

// s_waitcnt vmcnt(0): all of this wave's memory operations have completed
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

template <typename T>
__device__ __forceinline__ void pub_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T pub_load(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// v with its two lowest mantissa bits replaced by tag & 3 (a publication tag)
__device__ __forceinline__ double tagged(double v, int tag) {
  const long long b = __double_as_longlong(v);
  return __longlong_as_double((b & ~3LL) | (long long)(tag & 3));
}

// v + v of the lane that DPP control CTRL selects within each row of 16 lanes
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return v + __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// value of lane l (wave-uniform)
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)b, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Sum over the 64 lanes of a wave (every lane active), wave-uniform and in
// one fixed order: butterflies within each row of 16 lanes through DPP lane
// selects (lane ^ 1, lane ^ 2, 8-lane mirror, 16-lane mirror; a few cycles
// each, against an LDS round trip per step for a shuffle), then the four row
// sums read out of lanes 0, 16, 32, 48.
__device__ __forceinline__ double wave_sum(double v) {
  v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141>(v);  // row_half_mirror
  v = dpp_add<0x140>(v);  // row_mirror
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// workgroup state of the grid reductions (one set per kernel)
struct ReduceLds {
  double sm[kWaves][4];
  double tot[kWaves][4];
  int abort_flag;
};
__device__ __forceinline__ ReduceLds& reduce_lds() {
  __shared__ ReduceLds l;
  return l;
}

// First half of a grid reduction: the workgroup's partials go out.  With
// barrier = true every wave first waits for its earlier stores to complete,
// so that the reduction also orders them before every workgroup's later
// reads; without it the reduction carries values only.
template <int NV>
__device__ __forceinline__ void grid_post(GridSync& gs, const double (&v)[NV], bool barrier) {
  static_assert(NV >= 1 && NV <= 4, "at most 4 values per reduction");
  ReduceLds& L = reduce_lds();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) L.abort_flag = 0;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double w = wave_sum(v[i]);
    if (lane == 0) L.sm[wave][i] = w;
  }
  if (barrier) wait_vmem();  // this wave's published data has reached the coherence point
  __syncthreads();
  gs.epoch += 1u;
  const uint32_t ep = gs.epoch;
  const int slot = ep & 1u;
  // Partials are published as 64-bit words (epoch << 32 | one half of the
  // FP64 value): the value arrives together with the arrival signal, so
  // readers poll the partials themselves (no separate flag store, no
  // second round trip to read them).  A word is single-copy atomic, so a
  // matching tag means a matching half.  A slot is rewritten only two
  // epochs later, after every workgroup has passed the next barrier, so the
  // tag of the slot being read is either ep - 2 (not yet written) or ep.
  uint64_t* words = reinterpret_cast<uint64_t*>(gs.red) + gs.region;
  if (threadIdx.x < 2 * NV) {
    const int i = threadIdx.x >> 1;
    double acc = L.sm[0][i];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) acc += L.sm[w][i];
    const uint64_t bits = (uint64_t)__double_as_longlong(acc);
    const uint32_t half = (threadIdx.x & 1) ? (uint32_t)(bits >> 32) : (uint32_t)bits;
    pub_store(&words[((size_t)slot * MFEM_RED_MAX_BLOCKS + blockIdx.x) * 8 + threadIdx.x],
              ((uint64_t)ep << 32) | half);
  }
}

// Second half: wait for every workgroup's partials of the last grid_post and
// sum them.  Work placed between the two halves overlaps the round trip.
template <int NV>
__device__ __forceinline__ bool grid_wait(GridSync& gs, double (&v)[NV]) {
  ReduceLds& L = reduce_lds();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t ep = gs.epoch;
  const int slot = ep & 1u;
  const uint64_t* words = reinterpret_cast<const uint64_t*>(gs.red) + gs.region;
  // Reading: eight lanes per partial, lane j loading word j, so that one
  // load instruction fetches eight partials' lines whole (one request per
  // line instead of one per word: every workgroup polls every partial, and
  // the request count, not the latency alone, sets the cost of a poll).
  // Each lane polls its own word until its tag is the current epoch; the
  // even lane of a pair then takes its neighbour's half (DPP) and owns
  // value j / 2.  Up to kPollPasses partials per lane per batch of loads.
  const int j = threadIdx.x & 7;
  const int grp = threadIdx.x >> 3;  // 32 partials per pass of the block
  double a = 0.0;                    // value j / 2 of this lane's partials, in workgroup order
  bool to = false;
  for (int base = gs.wg0; base < gs.wg1; base += 32 * kPollPasses) {
    uint64_t wv[kPollPasses];
#pragma unroll
    for (int p = 0; p < kPollPasses; ++p) {
      const int b = base + grp + 32 * p;
      // (words not read count as arrived)
      wv[p] = (b < gs.wg1 && j < 2 * NV) ? pub_load(&words[((size_t)slot * MFEM_RED_MAX_BLOCKS + b) * 8 + j])
                                          : ((uint64_t)ep << 32);
    }
    // poll until every word of this lane carries the current epoch,
    // reloading the stale ones together (one round trip per poll)
    auto pending = [&]() {
      bool any = false;
#pragma unroll
      for (int p = 0; p < kPollPasses; ++p) any |= (uint32_t)(wv[p] >> 32) != ep;
      return any;
    };
    if (pending() && !to) {
      const uint64_t t0 = wall_clock64();
      do {
        // error bits and the time limit are looked at only once a wait is
        // long: the status load would add a round trip to every poll
        const uint64_t dt = wall_clock64() - t0;
        if (dt > kStatusCheckAfter && (dt > kBarrierTimeout || pub_load(&gs.status[MFEM_STATUS_ERRBITS]) != 0)) {
          to = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int p = 0; p < kPollPasses; ++p) {
          if ((uint32_t)(wv[p] >> 32) != ep)
            wv[p] = pub_load(&words[((size_t)slot * MFEM_RED_MAX_BLOCKS + base + grp + 32 * p) * 8 + j]);
        }
      } while (pending());
    }
    // (all lanes active again: DPP reads every lane of the quad)
#pragma unroll
    for (int p = 0; p < kPollPasses; ++p) {
      const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)wv[p], 0xB1, 0xF, 0xF, false);
      const bool mine = (j & 1) == 0 && j < 2 * NV && base + grp + 32 * p < gs.wg1;
      if (mine) a += __longlong_as_double((long long)(((uint64_t)hi << 32) | (uint32_t)wv[p]));
    }
  }
  if (to) {
    atomicOr(&gs.status[MFEM_STATUS_ERRBITS], 1);
    L.abort_flag = 1;
  }
  // per wave: value i sits in lanes 2 i + 8 g (g = 0..7); pairs of groups
  // one row apart through DPP (row_ror 8), then the four rows read out in
  // order
  a = dpp_add<0x128>(a);
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double w = (lane_value(a, 2 * i) + lane_value(a, 16 + 2 * i)) +
                     (lane_value(a, 32 + 2 * i) + lane_value(a, 48 + 2 * i));
    if (lane == 0) L.tot[wave][i] = w;
  }
  __syncthreads();
  if (L.abort_flag) {
    gs.abort = true;
    return false;
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double t = L.tot[0][i];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) t += L.tot[w][i];
    v[i] = lane_value(t, 0);  // (uniform: scalar registers)
  }
  // (no closing barrier: the next reduction writes tot only after its first
  // barrier, which every thread reaches after these reads; before it, it
  // writes sm, read here by no one, and clears abort_flag, which is clear
  // already, since a set flag ends the solve)
  return true;
}

template <int NV>
__device__ __forceinline__ bool grid_reduce(GridSync& gs, double (&v)[NV]) {
  grid_post(gs, v, true);
  return grid_wait(gs, v);
}

// ---------------------------------------------------------------------------
// persistent block-Jacobi PCG (warp.optim.linear.cg semantics: stop when the
// residual norm <= tol * |g|, at most cg_maxiter iterations)
//
// Each wave owns CPW chunks of 64 rows and keeps every per-row vector in
// registers for the whole solve.  With RW > 0 (one chunk per wave, chunk
// width <= RW) the wave's matrix rows (RW slots x 9 values + columns) are
// loaded into VGPRs once, so an iteration reads only the published vector.
// Published vectors ping-pong between two buffers (publish count parity).
//
// Standard variant (PIPE = false), two grid reductions per iteration:
//   1. y = H z; q = y + beta q (= H p); p = z + beta p; reduce p.q
//   2. x += alpha p; r -= alpha q; z = M r; publish z; reduce r.z, r.r
// Pipelined variant (PIPE = true; Ghysels & Vanroose 2014, "Hiding global
// synchronization latency in the preconditioned conjugate gradient
// algorithm", Alg. 4), one grid reduction per iteration that also publishes
// m = M w for the single SpMV n = H m.
// Both: when the recurrence residual passes (at tol/2), the true residual
// g - H x is recomputed with the FP64 iterate and FP64 accumulation; if it
// fails tol, the iteration restarts from it (residual replacement, <= 16).
// ---------------------------------------------------------------------------
template <typename T, typename A, int GRP = 8>
__device__ __forceinline__ void spmv_row_stream(const mfem_solver& S, int c, int lane, const T* vec, A* out,
                                                bool coherent) {
  A a0 = 0, a1 = 0, a2 = 0;
  const int s0 = S.chunk_ptr[c], s1 = S.chunk_ptr[c + 1];
  for (int sb = s0; sb < s1; sb += GRP) {
    int col[GRP];
    float vv[GRP][9];
    T g[GRP][3];
#pragma unroll
    for (int k = 0; k < GRP; ++k) {
      const bool ok = sb + k < s1;
      col[k] = ok ? S.sell_col[(sb + k) * 64 + lane] : 0;
      const float* vb = S.sell_val + (size_t)(ok ? sb + k : sb) * 576 + lane;
#pragma unroll
      for (int q = 0; q < 9; ++q) vv[k][q] = ok ? vb[64 * q] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < GRP; ++k)
#pragma unroll
      for (int a = 0; a < 3; ++a) g[k][a] = coherent ? pub_load(&vec[3 * col[k] + a]) : vec[3 * col[k] + a];
#pragma unroll
    for (int k = 0; k < GRP; ++k) {
      a0 += (A)vv[k][0] * g[k][0] + (A)vv[k][1] * g[k][1] + (A)vv[k][2] * g[k][2];
      a1 += (A)vv[k][3] * g[k][0] + (A)vv[k][4] * g[k][1] + (A)vv[k][5] * g[k][2];
      a2 += (A)vv[k][6] * g[k][0] + (A)vv[k][7] * g[k][1] + (A)vv[k][8] * g[k][2];
    }
  }
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
}

template <int RW, typename T, int GRP>
__device__ __forceinline__ void spmv_row_resident(const float (&val)[RW][9], const int (&col)[RW], const T* vec,
                                                  T* out) {
  T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int kb = 0; kb < RW; kb += GRP) {
    T g[GRP][3];  // gathers in flight per group
#pragma unroll
    for (int k = 0; k < GRP; ++k)
#pragma unroll
      for (int a = 0; a < 3; ++a) g[k][a] = pub_load(&vec[3 * col[kb + k] + a]);
#pragma unroll
    for (int k = 0; k < GRP; ++k) {
      const float* v = val[kb + k];
      a0 += (T)v[0] * g[k][0] + (T)v[1] * g[k][1] + (T)v[2] * g[k][2];
      a1 += (T)v[3] * g[k][0] + (T)v[4] * g[k][1] + (T)v[5] * g[k][2];
      a2 += (T)v[6] * g[k][0] + (T)v[7] * g[k][1] + (T)v[8] * g[k][2];
    }
  }
  if constexpr (sizeof(T) > sizeof(float)) {
    // FP64 products: materialise the sums here (sunk into the caller's
    // later branches they keep every gathered value and converted matrix
    // entry live at once, which spills)
    asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2));
  }
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
}

// resident-matrix storage; RW == 0 means "stream from HBM"
template <int RW>
struct RowBlocks {
  float val[RW][9];
  int col[RW];
};
template <>
struct RowBlocks<0> {};

template <int CPW, int RW, int GRP = 8>
__device__ __forceinline__ void spmv_rows(const mfem_solver& S, const int (&chunk)[CPW], const bool (&own)[CPW],
                                          int lane, const RowBlocks<RW>& rb, const float* vec, float (&y)[CPW][3]) {
  if constexpr (RW > 0) {
    static_assert(CPW == 1, "resident rows need one chunk per wave");
    spmv_row_resident<RW, float, 8>(rb.val, rb.col, vec, y[0]);
  } else {
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      y[k][0] = y[k][1] = y[k][2] = 0.f;
      if (own[k]) spmv_row_stream<float, float, GRP>(S, chunk[k], lane, vec, y[k], true);
    }
  }
}

// FP64 vector, FP64 accumulation: the true-residual checks and the pipelined
// variant's single SpMV per iteration (resident rows when available)
template <int CPW, int RW, bool RESIDENT>
__device__ __forceinline__ void spmv_rows64(const mfem_solver& S, const int (&chunk)[CPW], const bool (&own)[CPW],
                                            int lane, const RowBlocks<RW>& rb, const double* vec,
                                            double (&y)[CPW][3]) {
  if constexpr (RW > 0 && RESIDENT) {
    spmv_row_resident<RW, double, 2>(rb.val, rb.col, vec, y[0]);
  } else {
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      y[k][0] = y[k][1] = y[k][2] = 0.0;
      if (own[k]) spmv_row_stream<double, double, 2>(S, chunk[k], lane, vec, y[k], true);
    }
  }
}

// cache-policy bits of a direct-to-LDS load: SC1, device scope (the same
// coherence as the agent-scope loads of published vectors)
constexpr int kCachePolicyDevice = 16;
// 64-byte lines of a published FP64 vector staged per workgroup (32 KB of LDS)
constexpr int kHaloLines = 512;

// resident rows against the staged lines: off[k] is slot k's column's FP64
// offset in hv (3 consecutive values)
template <int RW>
__device__ __forceinline__ void spmv_row_lds(const float (&val)[RW][9], const int (&off)[RW], const double* hv,
                                             double* out) {
  constexpr int GRP = 4;  // LDS reads in flight per group (registers)
  double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
  for (int kb = 0; kb < RW; kb += GRP) {
#pragma unroll
    for (int k = kb; k < kb + GRP && k < RW; ++k) {
      const double* g = hv + off[k];
      const double g0 = g[0], g1 = g[1], g2 = g[2];
      // (opaque per use: the FP32 values are converted here, not hoisted out
      // of the solve loop as FP64 copies that overflow into AGPRs)
      float v[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        v[q] = val[k][q];
        asm volatile("" : "+v"(v[q]));
      }
      a0 += (double)v[0] * g0 + (double)v[1] * g1 + (double)v[2] * g2;
      a1 += (double)v[3] * g0 + (double)v[4] * g1 + (double)v[5] * g2;
      a2 += (double)v[6] * g0 + (double)v[7] * g1 + (double)v[8] * g2;
    }
    // sums materialised per group, reads not moved across (see
    // spmv_row_resident)
    asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2)::"memory");
  }
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
}

template <int CPW, typename T>
__device__ __forceinline__ void apply_M(const float (&M)[CPW][9], int k, const double* r, T* z) {
#pragma unroll
  for (int a = 0; a < 3; ++a) z[a] = (T)(M[k][3 * a] * r[0] + M[k][3 * a + 1] * r[1] + M[k][3 * a + 2] * r[2]);
}

// block-Jacobi inverse of a row's diagonal block (identity when the block is
// singular or preconditioning is off)
// (row < 0: identity)
__device__ __forceinline__ void block_jacobi(const mfem_solver& S, int row, float (&M)[9]) {
#pragma unroll
  for (int a = 0; a < 9; ++a) M[a] = (a % 4 == 0) ? 1.f : 0.f;
  if (S.precond && row >= 0) {
    float Dg[9];
    const int dl = S.diag_lane[row];
#pragma unroll
    for (int a = 0; a < 9; ++a) Dg[a] = S.sell_val[val_index(dl, a)];
    if (fabsf(det3(Dg)) >= S.sing_thr) inv3(Dg, M);
  }
}

// Sums over the workgroup of NV per-lane values, in one fixed order, so that
// every thread gets the same numbers: wave sums through DPP, then the four
// wave sums from LDS.  buf: [kWaves][24] doubles of LDS that no other thread
// reads until every thread has passed this call's barrier (callers alternate
// two buffers).  Every thread of the workgroup calls it together.
template <int NV>
__device__ __forceinline__ void wg_sum_post(const double (&v)[NV], double (*buf)[24]) {
  static_assert(NV <= 24, "at most 24 values");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double w = wave_sum(v[i]);
    if (lane == 0) buf[wave][i] = w;
  }
}
template <int NV>
__device__ __forceinline__ void wg_sum_get(double (&v)[NV], double (*buf)[24]) {
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = (buf[0][i] + buf[1][i]) + (buf[2][i] + buf[3][i]);
}
// (the two halves back to back; work placed between them overlaps the
// other waves' arrival at the barrier)
template <int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double (*buf)[24]) {
  wg_sum_post(v, buf);
  wg_sum_get(v, buf);
}

// rigid modes of a vertex at e (position about the aggregate's centroid over
// its radius): the 3 x 6 block [I | e_k x e], columns 3 + k = rotation about
// axis k
__device__ __forceinline__ void rigid_basis(float e0, float e1, float e2, float (&P)[3][6]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int j = 0; j < 6; ++j) P[a][j] = (j == a) ? 1.f : 0.f;
  P[0][4] = e2;
  P[0][5] = -e1;
  P[1][3] = -e2;
  P[1][5] = e0;
  P[2][3] = e1;
  P[2][4] = -e0;
}

// LOWREG (standard mode, streamed rows, one chunk per wave): the
// preconditioner blocks and the right-hand side are re-read from memory where
// used instead of held in registers and the phase clock is off, so that four
// waves fit per SIMD (twice the co-resident rows of the register-held form,
// for batches of many bodies).
template <int CPW, int RW, int MODE, bool LOWREG>
__device__ __forceinline__ void pcg_body(const mfem_solver& S, int newton_iter) {
  static_assert(!LOWREG || (CPW == 1 && RW == 0 && MODE == 0), "LOWREG: standard, streamed, one chunk per wave");
  constexpr bool PIPE = MODE != 0;
  constexpr bool PUB32 = MODE == 2;  // pipelined: publish the SpMV input in FP32
  // pipelined, resident rows, block Jacobi plus the workgroup-local
  // rigid-mode correction (the workgroup's 256 rows are one aggregate)
  constexpr bool COARSE = MODE == 3;
  static_assert(!COARSE || (CPW == 1 && RW > 0), "the coarse correction needs one chunk per wave and resident rows");
  // pipelined with resident rows: the FP64 SpMV input is staged in LDS, and
  // every FP64 publication carries its publication count in the two lowest
  // mantissa bits (a relative change below 1e-15)
  constexpr bool HALO = PIPE && !PUB32 && RW > 0;
  // a barrier timed out earlier: flags and epoch are inconsistent until the
  // host clears status and flags, so do nothing
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;
  // one body: all workgroups, chunks strided over the grid; a batch: the
  // body's workgroups, its own epoch and stats (workgroup w covers the
  // 256-row blocks [w*CPW, (w+1)*CPW); bodies start at multiples of CPW
  // blocks, so every workgroup belongs to one body)
  const bool batch = S.n_bodies > 1;
  const int body = batch ? body_of(S.body_wg, S.n_bodies, (int)blockIdx.x * CPW) : 0;
  const int wg0 = batch ? S.body_wg[body] / CPW : 0;
  const int wg1 = batch ? S.body_wg[body + 1] / CPW : (int)gridDim.x;
  int32_t* ep_ptr = batch ? S.body_epoch + body : S.status + MFEM_STATUS_EPOCH;
  GridSync gs{S.red, S.flags, S.status, (uint32_t)*ep_ptr, false, wg0, wg1, 0};
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const int nw = gridDim.x * kWaves;
  const size_t nrow_pad = (size_t)S.n_chunks * 64;
  int chunk[CPW];
  bool live[CPW], own[CPW];
  double x[CPW][3], r[CPW][3];
  float M[CPW][9], b[CPW][3];
  // standard variant: p, q, z in FP32
  float p[CPW][3], q[CPW][3], z[CPW][3];
  // pipelined variant: w, p, s, z in FP64; u = M r is recomputed where used
  // (u at a restart and u, q of the FP32-published form are held)
  double u[CPW][3], w[CPW][3], pd[CPW][3], sd[CPW][3], qd[CPW][3], zd[CPW][3];
  RowBlocks<RW> rb;
  int pub = 0, pubx = 0;  // publish counters: buffer = count & 1
  auto pzbuf = [&](int k) { return S.pz + (size_t)(k & 1) * nrow_pad * 3; };
  auto pxbuf = [&](int k) { return S.px + (size_t)(k & 1) * nrow_pad * 3; };
  // FP64 publication number k (buffer k & 1): the value with tag k & 3
  auto tagx = [&](double v, int k) -> double {
    if constexpr (HALO) return tagged(v, k);
    else return v;
  };

  double bb = 0.0;
  double* px0 = pxbuf(pubx);
#pragma unroll
  for (int k = 0; k < CPW; ++k) {
    chunk[k] = batch ? ((int)blockIdx.x * CPW + k) * kWaves + (int)(threadIdx.x >> 6) : gw + k * nw;
    own[k] = chunk[k] < S.n_chunks;
    const int row = chunk[k] * 64 + lane;
    live[k] = own[k] && row < S.n_verts;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      x[k][a] = r[k][a] = u[k][a] = w[k][a] = 0.0;
      pd[k][a] = sd[k][a] = qd[k][a] = zd[k][a] = 0.0;
      b[k][a] = 0.f;
      p[k][a] = q[k][a] = z[k][a] = 0.f;
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) M[k][a] = (a % 4 == 0) ? 1.f : 0.f;
    if (live[k]) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        b[k][a] = S.g[3 * row + a];
        x[k][a] = (double)S.x_tilde[3 * row + a] - (double)S.x[3 * row + a];
      }
      if constexpr (!LOWREG) block_jacobi(S, row, M[k]);
      bb += (double)b[k][0] * b[k][0] + (double)b[k][1] * b[k][1] + (double)b[k][2] * b[k][2];
    }
    if (own[k]) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        pub_store(&px0[3 * row + a], tagx(x[k][a], 0));
        // the other buffer as publication -1 (tag 3): whatever an earlier
        // launch left there cannot carry the tag of this launch's first
        // publication into it
        if constexpr (HALO) pub_store(&pxbuf(1)[3 * row + a], tagx(x[k][a], 3));
      }
    }
  }
  ++pubx;
  if constexpr (RW > 0) {
    // this wave's matrix rows, resident for the whole solve (padding: 0, col 0)
    const int c = chunk[0];
    const int s0 = own[0] ? S.chunk_ptr[c] : 0;
    const int wd = own[0] ? S.chunk_ptr[c + 1] - s0 : 0;
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const bool ok = k < wd;
      rb.col[k] = ok ? S.sell_col[(s0 + k) * 64 + lane] : 0;
      const float* vb = S.sell_val + (size_t)(ok ? s0 + k : 0) * 576 + lane;
#pragma unroll
      for (int a = 0; a < 9; ++a) rb.val[k][a] = ok ? vb[64 * a] : 0.f;
    }
  }
  // device-clock profile of the solve phases (wave 0 of workgroup 0 only)
  const bool prof = !LOWREG && blockIdx.x == wg0 && threadIdx.x < 64;
  // LOWREG: right-hand side and preconditioner re-read where used
  auto reload_b = [&]() {
    if constexpr (LOWREG) {
#pragma unroll
      for (int k = 0; k < CPW; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) b[k][a] = live[k] ? S.g[3 * (chunk[k] * 64 + lane) + a] : 0.f;
    }
  };
  auto reload_M = [&](int k) {
    if constexpr (LOWREG) block_jacobi(S, live[k] ? chunk[k] * 64 + lane : -1, M[k]);
  };
  uint64_t tk[6] = {0, 0, 0, 0, 0, 0};
  uint64_t tprev = wall_clock64();
#define PCG_LAP(k)                        \
  if (prof) {                             \
    const uint64_t tnow = wall_clock64(); \
    tk[k] += tnow - tprev;                \
    tprev = tnow;                         \
  }
  // Pipelined, resident rows: the FP64 SpMV input is staged per workgroup.
  // The workgroup's columns touch a few hundred distinct 64-byte lines of the
  // published vector (each 24-byte row spans one or two); gathering them lane
  // by lane costs ~25 line lookups per wave instruction, 48 instructions per
  // wave.  Instead the lines are copied once per SpMV into LDS with
  // coalesced direct-to-LDS loads (16 bytes a lane, device scope) and the
  // gathers read LDS.  Set-up, once per launch: mark the lines in an LDS
  // bitmap, count them (exclusive prefix over the bitmap words), list them in
  // ascending order (a row that spans two lines finds them adjacent) and turn
  // each resident slot's column into its LDS offset.  Slot 0 is a line of
  // zeros for the padding slots.  A workgroup with more lines than fit
  // streams its rows from memory instead.
  constexpr int kHaloQ = kHaloLines * 4 / kBlock;  // 16-byte quarters per thread
  __shared__ double hv[HALO ? kHaloLines * 8 : 1];
  // hl[1 + j]: line of staged slot j + 1, -1 past the last one
  __shared__ int hl[HALO ? kHaloLines + 1 : 1];
  __shared__ int hcount;
  bool staged = false;
  // per-wave transpose slice for the coalesced publication of m
  __shared__ double tr[(PIPE && !PUB32) ? kWaves * 192 : 1];
  double* trw = tr + ((PIPE && !PUB32) ? (threadIdx.x >> 6) * 192 : 0);
  // Workgroup-local coarse correction, set up once per launch (before the
  // halo set-up, which rewrites the resident columns and reuses hv):
  //   P_i   = rigid modes of row i about the aggregate's centroid,
  //   P'_i  = P_i - w_i D_i^-1 (H_aa P)_i  (H_aa: columns of this workgroup),
  //   Ac    = P'^T H_aa P' (6 x 6, FP64), cinv = its pseudo-inverse (pivoted
  //           Cholesky, pivots <= 1e-9 of the diagonal dropped),
  //   M r   = D^-1 r + P' cinv P'^T r.
  // Pc holds this row's P' for the whole solve; the coarse coefficients
  // cinv P'^T v of w, r and s are wave-uniform 6-vectors (ccw, ccr, ccs).
  float Pc[3][6];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int j = 0; j < 6; ++j) Pc[a][j] = 0.f;
  double Pd[3][6];  // (Pc in FP64 for the solve: no conversion per use)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int j = 0; j < 6; ++j) Pd[a][j] = 0.0;
  __shared__ double cinv[COARSE ? 21 : 1];
  __shared__ double cred[COARSE ? 2 : 1][kWaves][24];
  int cpar = 0;  // which cred buffer the next workgroup sum uses
  double ccw[6], ccr[6], ccs[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) ccw[j] = ccr[j] = ccs[j] = 0.0;
  if constexpr (COARSE) {
    const int row = chunk[0] * 64 + lane;
    const bool lv = live[0];
    const int li = (int)threadIdx.x;  // this row's place in the aggregate (one chunk per wave)
    // (the position and the smoothing weight loaded together)
    float xi[3] = {0.f, 0.f, 0.f};
    if (lv) {
#pragma unroll
      for (int a = 0; a < 3; ++a) xi[a] = S.x[3 * row + a];
    }
    const float om = lv ? (S.coarse_w ? S.coarse_w[row] : S.coarse_omega) : 0.f;
    // centroid and RMS radius from one workgroup sum: sum |x - c|^2 =
    // sum |x|^2 - |sum x|^2 / n (FP64 over FP32 positions; the radius only
    // scales the rotation modes, which leaves M unchanged)
    double m5[5] = {xi[0], xi[1], xi[2], (double)xi[0] * xi[0] + (double)xi[1] * xi[1] + (double)xi[2] * xi[2],
                    lv ? 1.0 : 0.0};
    wg_sum<5>(m5, cred[cpar]);
    cpar ^= 1;
    const double cnt = m5[4];
    double d[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) d[a] = (lv && cnt > 0.0) ? (double)xi[a] - m5[a] / cnt : 0.0;
    const double r2 = cnt > 0.0 ? m5[3] - (m5[0] * m5[0] + m5[1] * m5[1] + m5[2] * m5[2]) / cnt : 0.0;
    const double rad = (cnt > 0.0 && r2 > 0.0) ? sqrt(r2 / cnt) : 1.0;
    const float e0 = (float)(d[0] / rad), e1 = (float)(d[1] / rad), e2 = (float)(d[2] / rad);
    if (lv) rigid_basis(e0, e1, e2, Pc);
    // the aggregate's positions, then its smoothed basis, staged in hv
    float4* els = reinterpret_cast<float4*>(hv);
    float* pls = reinterpret_cast<float*>(hv) + 4 * kBlock;
    static_assert(4 * kBlock + 18 * kBlock <= 2 * kHaloLines * 8, "coarse set-up scratch fits the halo buffer");
    els[li] = make_float4(e0, e1, e2, lv ? 1.f : 0.f);
    // resident slot k's column inside the aggregate, or -1
    const int cs0 = own[0] ? S.chunk_ptr[chunk[0]] : 0;
    const int cwd = own[0] ? S.chunk_ptr[chunk[0] + 1] - cs0 : 0;
    int lpos[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int c = rb.col[k];
      lpos[k] = (lv && k < cwd && (c >> 8) == (int)blockIdx.x) ? (c & 255) : -1;
    }
    __syncthreads();
    if (om != 0.f) {
      float HP[3][6];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 6; ++j) HP[a][j] = 0.f;
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        if (lpos[k] < 0) continue;
        const float4 e = els[lpos[k]];
        float Pj[3][6];
        rigid_basis(e.x, e.y, e.z, Pj);
        const float* v = rb.val[k];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int j = 0; j < 6; ++j)
            HP[a][j] += e.w * (v[3 * a] * Pj[0][j] + v[3 * a + 1] * Pj[1][j] + v[3 * a + 2] * Pj[2][j]);
      }
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 6; ++j)
          Pc[a][j] -= om * (M[0][3 * a] * HP[0][j] + M[0][3 * a + 1] * HP[1][j] + M[0][3 * a + 2] * HP[2][j]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int j = 0; j < 6; ++j) pls[li * 18 + 6 * a + j] = Pc[a][j];
    __syncthreads();
    float HQ[3][6];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int j = 0; j < 6; ++j) HQ[a][j] = 0.f;
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      if (lpos[k] < 0) continue;
      const float* q = pls + lpos[k] * 18;
      const float* v = rb.val[k];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 6; ++j) HQ[a][j] += v[3 * a] * q[j] + v[3 * a + 1] * q[6 + j] + v[3 * a + 2] * q[12 + j];
    }
    double ac[21];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = i; j < 6; ++j)
        ac[sym21(i, j)] = (double)Pc[0][i] * HQ[0][j] + (double)Pc[1][i] * HQ[1][j] + (double)Pc[2][i] * HQ[2][j];
    wg_sum<21>(ac, cred[cpar]);
    cpar ^= 1;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int j = 0; j < 6; ++j) Pd[a][j] = (double)Pc[a][j];
    if (threadIdx.x == 0) {
      double A[6][6], L[6][6], Li[6][6];
      bool kept[6];
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          A[i][j] = ac[sym21(i, j)];
          L[i][j] = 0.0;
        }
      // (one division per pivot: rl = 1 / L_jj, products elsewhere)
      double rl[6];
      for (int j = 0; j < 6; ++j) {
        double dj = A[j][j];
        for (int q = 0; q < j; ++q) dj -= L[j][q] * L[j][q];
        kept[j] = A[j][j] > 0.0 && dj > 1e-9 * A[j][j];
        L[j][j] = kept[j] ? sqrt(dj) : 1.0;
        rl[j] = 1.0 / L[j][j];
        for (int i = j + 1; i < 6; ++i) {
          double v = A[j][i];
          for (int q = 0; q < j; ++q) v -= L[i][q] * L[j][q];
          L[i][j] = kept[j] ? v * rl[j] : 0.0;
        }
      }
      for (int i = 0; i < 6; ++i)  // a dropped row takes no part either
        if (!kept[i])
          for (int q = 0; q < i; ++q) L[i][q] = 0.0;
      for (int c = 0; c < 6; ++c)
        for (int i = 0; i < 6; ++i) {
          double v = (i == c) ? 1.0 : 0.0;
          for (int q = 0; q < i; ++q) v -= L[i][q] * Li[q][c];
          Li[i][c] = v * rl[i];
        }
      for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
          double v = 0.0;
          if (kept[i] && kept[j])
            for (int q = 0; q < 6; ++q) v += kept[q] ? Li[q][i] * Li[q][j] : 0.0;
          cinv[sym21(i, j)] = v;
        }
    }
    __syncthreads();
  }
  // coarse coefficients cinv P'^T v of a per-row vector v (all threads
  // together), and M v = D^-1 v + P' y with such coefficients y
  // (in two halves: coarse_post sends this lane's share P'^T v, coarse_get
  // waits for the workgroup and forms cinv P'^T v)
  auto coarse_post = [&](const double (&v)[3]) {
    double t[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) t[j] = Pd[0][j] * v[0] + Pd[1][j] * v[1] + Pd[2][j] * v[2];
    wg_sum_post<6>(t, cred[cpar]);
  };
  auto coarse_get = [&](double (&y)[6]) {
    double t[6];
    wg_sum_get<6>(t, cred[cpar]);
    cpar ^= 1;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) acc += cinv[sym21(i, j)] * t[j];
      y[i] = acc;
    }
  };
  auto coarse_coef = [&](const double (&v)[3], double (&y)[6]) {
    coarse_post(v);
    coarse_get(y);
  };
  auto apply_Mc = [&](const double (&v)[3], const double (&y)[6], double (&out)[3]) {
    apply_M<CPW, double>(M, 0, v, out);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      out[a] += Pd[a][0] * y[0] + Pd[a][1] * y[1] + Pd[a][2] * y[2] + Pd[a][3] * y[3] + Pd[a][4] * y[4] +
                Pd[a][5] * y[5];
  };
  (void)coarse_coef;
  (void)coarse_post;
  (void)coarse_get;
  (void)apply_Mc;

  if constexpr (HALO) {
    uint32_t* bm = reinterpret_cast<uint32_t*>(hv);
    const long long n_lines = (long long)nrow_pad * 3 / 8;
    const int nwords = (int)((n_lines + 31) >> 5);
    // bitmap and prefix share the line buffer (dead once the offsets are set)
    if (2LL * nwords <= (long long)kHaloLines * 16) {
      uint32_t* pre = bm + nwords;
      const int c0 = own[0] ? S.chunk_ptr[chunk[0]] : 0;
      const int wd = own[0] ? S.chunk_ptr[chunk[0] + 1] - c0 : 0;
      for (int i = threadIdx.x; i < nwords; i += kBlock) bm[i] = 0u;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        if (k < wd) {
          const long long b0 = 24LL * rb.col[k];
          const int l0 = (int)(b0 >> 6), l1 = (int)((b0 + 23) >> 6);
          atomicOr(&bm[l0 >> 5], 1u << (l0 & 31));
          if (l1 != l0) atomicOr(&bm[l1 >> 5], 1u << (l1 & 31));
        }
      }
      __syncthreads();
      // exclusive prefix of the words' bit counts: contiguous word segments
      // per thread, then a scan of the 256 segment totals
      const int seg = (nwords + kBlock - 1) / kBlock;
      const int w0 = min((int)threadIdx.x * seg, nwords), w1 = min(w0 + seg, nwords);
      int mine = 0;
      for (int i = w0; i < w1; ++i) mine += __popc(bm[i]);
      int incl = mine;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
      }
      if (lane == 63) hl[threadIdx.x >> 6] = incl;  // wave totals (hl reused)
      __syncthreads();
      int base = 0;
      for (int wv = 0; wv < (int)(threadIdx.x >> 6); ++wv) base += hl[wv];
      if (threadIdx.x == kBlock - 1) hcount = base + incl;
      __syncthreads();
      const int H = hcount;
      staged = H <= kHaloLines - 1;
      if (staged) {
        for (int i = 1 + H + (int)threadIdx.x; i <= kHaloLines; i += kBlock) hl[i] = -1;
        int run = base + incl - mine;
        for (int i = w0; i < w1; ++i) {
          pre[i] = (uint32_t)run;
          uint32_t bits = bm[i];
          while (bits) {
            const int j = __ffs(bits) - 1;
            bits &= bits - 1u;
            hl[1 + run++] = i * 32 + j;
          }
        }
      }
      __syncthreads();
      if (staged) {
#pragma unroll
        for (int k = 0; k < RW; ++k) {
          int off = 0;
          if (k < wd) {
            const int c = rb.col[k];
            const int l0 = (int)((24LL * c) >> 6);
            const uint32_t wbits = bm[l0 >> 5] & ((1u << (l0 & 31)) - 1u);
            off = (1 + (int)pre[l0 >> 5] + __popc(wbits)) * 8 + ((3 * c) & 7);
          }
          rb.col[k] = off;
        }
      }
      __syncthreads();
      if (staged && threadIdx.x < 8) hv[threadIdx.x] = 0.0;
      // (the first staging comes after a grid reduction, whose workgroup
      // barrier orders these writes before it)
    }
  }
  // (a workgroup with too many lines for its buffer streams its SpMV input
  // and needs the barrier before every SpMV: counted here)
  double red1[2] = {bb, (HALO && !staged && threadIdx.x == 0) ? 1.0 : 0.0};
  if (!grid_reduce(gs, red1)) return;
  const double bnorm2 = red1[0];
  // pipelined, every workgroup staged: the main loop's SpMV overlaps its
  // grid reduction and checks the publication tags of what it stages
  const bool ovl = HALO && red1[1] == 0.0;
  const double tol2 = (double)S.cg_tol * (double)S.cg_tol * bnorm2;
  // the recurrence runs to a tighter residual so that the FP64 re-check
  // normally passes without a restart: 0.5 tol with FP32 directions
  // (standard), 0.9 tol for the pipelined form, whose FP64 recurrences
  // track the true residual closely
  const double tol2_rec = (PIPE ? 0.81 : 0.25) * tol2;
  int it = 0, restarts = 0;
  double rt2 = 0.0;

  // Standard mode: z = M r for the wave's chunks.  Block Jacobi, then (when
  // local_sweeps > 0) damped block-Jacobi sweeps on the workgroup's own rows,
  // z <- z + omega D^-1 (r - A_loc z), with the workgroup's z exchanged in
  // LDS: A_loc keeps only the columns whose rows this workgroup owns.
  // alternate buffers (one barrier per sweep); slot kZero of each holds zeros
  // for the columns of other workgroups
  constexpr int kZero = CPW * kWaves * 64;
  __shared__ float4 zl[2][kZero + 1];
  if (threadIdx.x < 2) zl[threadIdx.x][kZero] = make_float4(0.f, 0.f, 0.f, 0.f);
  const int wave = threadIdx.x >> 6;
  // LDS slot of column vertex c, or -1 when another workgroup owns it (one
  // chunk per wave: workgroup w owns rows [256 w, 256 w + 256) in both
  // layouts)
  auto local_pos = [&](int c) -> int {
    if constexpr (CPW == 1) {
      return (c >> 8) == (int)blockIdx.x ? (c & 255) : -1;
    } else {
      const int cc = c >> 6;
      int li;
      if (batch) {
        li = cc - (int)blockIdx.x * kWaves * CPW;
        if (li < 0 || li >= kWaves * CPW) return -1;
      } else {
        const int g = cc % nw;
        if (g / kWaves != (int)blockIdx.x) return -1;
        li = (cc / nw) * kWaves + g % kWaves;
      }
      return li * 64 + (c & 63);
    }
  };
  // resident rows: each slot's LDS slot for the sweeps, found once
  int lp[RW > 0 ? RW : 1];
  if constexpr (RW > 0) {
#pragma unroll
    for (int kk = 0; kk < RW; ++kk) {
      const int pos = local_pos(rb.col[kk]);
      lp[kk] = pos < 0 ? kZero : pos;
    }
  }
  auto precondition = [&]() {
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
      if (!own[k]) continue;
      reload_M(k);
      apply_M<CPW, float>(M, k, r[k], z[k]);
    }
    const int ns = S.local_sweeps;
    const float om = S.local_omega;
    for (int sw = 0; sw < ns; ++sw) {
      // (sweep sw writes buffer sw & 1; a wave can only rewrite it two
      // sweeps later, after every wave has passed the next barrier)
      float4* zb4 = zl[sw & 1];
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        if (!own[k]) continue;
        zb4[(k * kWaves + wave) * 64 + lane] = make_float4(z[k][0], z[k][1], z[k][2], 0.f);
      }
      __syncthreads();
      float t[CPW][3];
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        auto add = [&](int c, const float* v) {
          const int pos = local_pos(c);
          if (pos < 0) return;
          const float4 zj = zb4[pos];
          const float z0 = zj.x, z1 = zj.y, z2 = zj.z;
          a0 += v[0] * z0 + v[1] * z1 + v[2] * z2;
          a1 += v[3] * z0 + v[4] * z1 + v[5] * z2;
          a2 += v[6] * z0 + v[7] * z1 + v[8] * z2;
        };
        if (own[k]) {
          if constexpr (RW > 0) {
#pragma unroll
            for (int kk = 0; kk < RW; ++kk) {
              const float4 zj = zb4[lp[kk]];
              const float* v = rb.val[kk];
              a0 += v[0] * zj.x + v[1] * zj.y + v[2] * zj.z;
              a1 += v[3] * zj.x + v[4] * zj.y + v[5] * zj.z;
              a2 += v[6] * zj.x + v[7] * zj.y + v[8] * zj.z;
            }
          } else {
            for (int sl = S.chunk_ptr[chunk[k]]; sl < S.chunk_ptr[chunk[k] + 1]; ++sl) {
              float v[9];
              const float* vb = S.sell_val + (size_t)sl * 576 + lane;
#pragma unroll
              for (int q = 0; q < 9; ++q) v[q] = vb[64 * q];
              add(S.sell_col[sl * 64 + lane], v);
            }
          }
        }
        t[k][0] = (float)r[k][0] - a0;
        t[k][1] = (float)r[k][1] - a1;
        t[k][2] = (float)r[k][2] - a2;
      }
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        if (!own[k]) continue;
        reload_M(k);
#pragma unroll
        for (int a = 0; a < 3; ++a)
          z[k][a] += om * (M[k][3 * a] * t[k][0] + M[k][3 * a + 1] * t[k][1] + M[k][3 * a + 2] * t[k][2]);
      }
    }
    // (the next call's first writes come after a grid reduction, whose
    // workgroup barrier ends this call's reads)
  };

  // y = H vec on the resident rows, FP64 products (pipelined variant)
  auto spmv_resident64 = [&](const double* vec, double (&y)[CPW][3]) {
    if constexpr (HALO) {
      if (staged) {
        int hline[kHaloQ];
#pragma unroll
        for (int i = 0; i < kHaloQ; ++i) hline[i] = hl[1 + ((threadIdx.x + kBlock * i) >> 2)];
#pragma unroll
        for (int i = 0; i < kHaloQ; ++i) {
          if (hline[i] >= 0) {
            const int q = threadIdx.x + kBlock * i;
            // lane l of the wave lands at base + 16 l: quarter q at hv[8 + 2 q]
            __builtin_amdgcn_global_load_lds(vec + (size_t)hline[i] * 8 + (q & 3) * 2,
                                             (__attribute__((address_space(3))) void*)&hv[8 + (q & ~63) * 2], 16,
                                             0, kCachePolicyDevice);
          }
        }
        wait_vmem();
        __syncthreads();
        spmv_row_lds<RW>(rb.val, rb.col, hv, y[0]);
      } else {
        // (too many lines for the buffer: rows streamed from memory)
        spmv_rows64<CPW, RW, false>(S, chunk, own, lane, rb, vec, y);
      }
    } else {
      spmv_rows64<CPW, RW, true>(S, chunk, own, lane, rb, vec, y);
    }
  };
  // The main loop's SpMV without a barrier before it: vec is publication
  // number tag (mod 4).  The lines are staged at once; each lane then checks
  // the tags of the two values of each quarter it loaded and the wave loads
  // again every quarter that some lane found stale (a neighbour's stores
  // still in flight), until all carry the tag.  A buffer is rewritten two
  // publications later, after a grid reduction that every workgroup enters
  // only after its SpMV of this one, so a stale value carries tag - 2.
  // The wait is bounded like a grid reduction's.
  __shared__ int stage_abort;
  if (threadIdx.x == 0) stage_abort = 0;  // (ordered by the next workgroup barrier)
  auto spmv_tagged = [&](const double* vec, int tag, double (&y)[CPW][3]) -> bool {
    if constexpr (HALO) {
      int hline[kHaloQ];
#pragma unroll
      for (int i = 0; i < kHaloQ; ++i) hline[i] = hl[1 + ((threadIdx.x + kBlock * i) >> 2)];
      auto load_q = [&](int i) {
        const int q = threadIdx.x + kBlock * i;
        __builtin_amdgcn_global_load_lds(vec + (size_t)hline[i] * 8 + (q & 3) * 2,
                                         (__attribute__((address_space(3))) void*)&hv[8 + (q & ~63) * 2], 16, 0,
                                         kCachePolicyDevice);
      };
#pragma unroll
      for (int i = 0; i < kHaloQ; ++i)
        if (hline[i] >= 0) load_q(i);
      wait_vmem();
      // bit i: quarter i holds a value without the tag.  All quarters are
      // read, without branches, so that the LDS reads go out together (a
      // quarter past the staged lines reads quarter 0 and is ignored).
      auto stale_bits = [&]() {
        uint32_t lo[kHaloQ][2];
#pragma unroll
        for (int i = 0; i < kHaloQ; ++i) {
          const int q = hline[i] >= 0 ? (int)threadIdx.x + kBlock * i : 0;
          const uint32_t* w = reinterpret_cast<const uint32_t*>(&hv[8 + 2 * q]);
          lo[i][0] = w[0];
          lo[i][1] = w[2];
        }
        uint32_t st = 0;
#pragma unroll
        for (int i = 0; i < kHaloQ; ++i) {
          const uint32_t bad = (((lo[i][0] ^ (uint32_t)tag) | (lo[i][1] ^ (uint32_t)tag)) & 3u) != 0u;
          st |= (bad & (uint32_t)(hline[i] >= 0)) << i;
        }
        return st;
      };
      uint32_t st = stale_bits();
      bool to = false;
      if (__ballot(st != 0)) {
        const uint64_t t0 = wall_clock64();
        do {
          const uint64_t dt = wall_clock64() - t0;
          if (dt > kStatusCheckAfter &&
              (dt > kBarrierTimeout || pub_load(&S.status[MFEM_STATUS_ERRBITS]) != 0)) {
            to = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
#pragma unroll
          for (int i = 0; i < kHaloQ; ++i)
            if (__ballot((st >> i) & 1u))  // (uniform: every lane of the quarter loads again)
              if (hline[i] >= 0) load_q(i);
          wait_vmem();
          st = stale_bits();
        } while (__ballot(st != 0));
      }
      if (to) {
        atomicOr(&S.status[MFEM_STATUS_ERRBITS], 1);
        stage_abort = 1;
      }
      __syncthreads();
      if (stage_abort) return false;
      spmv_row_lds<RW>(rb.val, rb.col, hv, y[0]);
      return true;
    } else {
      return false;
    }
  };
  // y = H vec in FP64 for the starts and the true-residual checks: staged
  // when the pipelined variant stages, else rows streamed from memory (the
  // resident FP64 form would hold too many registers in the standard one)
  auto spmv_check64 = [&](const double* vec, double (&y)[CPW][3]) {
    if constexpr (HALO) spmv_resident64(vec, y);
    else spmv_rows64<CPW, RW, false>(S, chunk, own, lane, rb, vec, y);
  };

  // (re)start from the current r.  Standard: z = M r published, (r.z, r.r)
  // reduced.  Pipelined: u = M r published (FP64), then w = H u (inside the
  // reduction's round trip when every workgroup stages its SpMV input).
  // rs_extra: one more per-lane value that the restart's reduction sums (the
  // Galerkin start's |g - H c_0|^2), 0 otherwise.
  double rs_extra = 0.0;
  auto restart = [&](double& rz, double& rr) -> bool {
    rz = 0.0;
    rr = 0.0;
    if constexpr (PUB32) {
      float* ub = pzbuf(pub);
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        if (!own[k]) continue;
        const int row = chunk[k] * 64 + lane;
        apply_M<CPW, double>(M, k, r[k], u[k]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          pub_store(&ub[3 * row + a], (float)u[k][a]);
          rr += r[k][a] * r[k][a];
        }
      }
      ++pub;
      double red2[2] = {rr, rs_extra};
      if (!grid_reduce(gs, red2)) return false;
      rr = red2[0];
      rs_extra = red2[1];
      float y[CPW][3];
      spmv_rows<CPW, RW>(S, chunk, own, lane, rb, ub, y);
#pragma unroll
      for (int k = 0; k < CPW; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) w[k][a] = own[k] ? (double)y[k][a] : 0.0;
    } else if constexpr (PIPE) {
      double* ub = pxbuf(pubx);
      if constexpr (COARSE) coarse_coef(r[0], ccr);
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        if (!own[k]) continue;
        const int row = chunk[k] * 64 + lane;
        if constexpr (COARSE) apply_Mc(r[k], ccr, u[k]);
        else apply_M<CPW, double>(M, k, r[k], u[k]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          pub_store(&ub[3 * row + a], tagx(u[k][a], pubx));
          rr += r[k][a] * r[k][a];
        }
      }
      ++pubx;
      double red2[2] = {rr, rs_extra};
      if (ovl) {
        grid_post(gs, red2, false);
        if (!spmv_tagged(ub, (pubx - 1) & 3, w)) return false;
        if (!grid_wait(gs, red2)) return false;
      } else {
        if (!grid_reduce(gs, red2)) return false;
        spmv_resident64(ub, w);
      }
      rr = red2[0];
      rs_extra = red2[1];
      if constexpr (COARSE) {
        coarse_coef(w[0], ccw);
#pragma unroll
        for (int j = 0; j < 6; ++j) ccs[j] = 0.0;
      }
    } else {
      float* zb = pzbuf(pub);
      precondition();
#pragma unroll
      for (int k = 0; k < CPW; ++k) {
        if (!own[k]) continue;
        const int row = chunk[k] * 64 + lane;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          pub_store(&zb[3 * row + a], z[k][a]);
          rz += r[k][a] * z[k][a];
          rr += r[k][a] * r[k][a];
        }
      }
      ++pub;
      double red2[3] = {rz, rr, rs_extra};
      if (!grid_reduce(gs, red2)) return false;
      rz = red2[0];
      rr = red2[1];
      rs_extra = red2[2];
    }
    return true;
  };

  int guess_used = 0;  // 0 x~ - x, 3 Galerkin start over the candidates
  double guess_ratio = 0.0;
  double guess_rrp = 0.0;  // |g - H x0|^2 of the plain start (Galerkin start only)
  if (bnorm2 > 0.0) {
    double rz = 0.0, rr = 0.0;
    {
      double y[CPW][3];
      spmv_check64(px0, y);
      reload_b();
#pragma unroll
      for (int k = 0; k < CPW; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) r[k][a] = live[k] ? (double)b[k][a] - y[k][a] : 0.0;
    }
    if ((S.pcg_flags & MFEM_PCG_LEARNED_GUESS) && S.dx_hist) {
      // Learned warm start, Galerkin form.  Six candidate starts c_0..c_5:
      // x~ - x (the plain guess), the learned guess, the previous Newton step,
      // the two before it, and this Newton iteration's solution in the
      // previous time step.  The start is the x0 in their span with the
      // lowest quadratic-model energy Phi(x0) = x0.H x0 / 2 - g.x0, i.e.
      // the smallest energy-norm error, the norm CG minimises: x0 = C a
      // with (C^T H C) a = C^T g.  It is never worse than the best single
      // candidate.  Gram entries from one SpMV per candidate (the histories
      // were written by the previous launch, so only the learned guess is
      // published), seven 4-value grid reductions (five, overlapped with the
      // SpMVs, without a learned guess), a 6 x 6 Cholesky that
      // drops candidates already in the span of the earlier ones (pivot <=
      // 1e-10 of the diagonal: equal or empty histories), then r0 = g - H x0.
      const double* h0 = S.dx_hist;
      const double* h1 = S.dx_hist + nrow_pad * 3;
      // c_5: this Newton iteration's solution in the previous time step
      const bool has5 = S.dx_same && newton_iter >= 0 && newton_iter < S.dx_same_iters;
      const double* h5 = has5 ? S.dx_same + (size_t)newton_iter * nrow_pad * 3 : nullptr;
      auto cget = [&](int i, int k, int a) -> double {
        const int row = chunk[k] * 64 + lane;
        if (!live[k]) return 0.0;
        if (i == 0) return x[k][a];
        if (i == 1) return S.guess ? (double)S.guess[3 * row + a] : 0.0;  // (no network: dropped as empty)
        if (i == 2) return S.dx[3 * row + a];
        if (i == 3) return h0[3 * row + a];
        if (i == 4) return h1[3 * row + a];
        return has5 ? h5[3 * row + a] : 0.0;
      };
      // sum over this lane's rows of c_i . y (y: an SpMV result) or c_i . g
      auto cdot = [&](int i, const double (&y)[CPW][3]) -> double {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < CPW; ++k)
#pragma unroll
          for (int a = 0; a < 3; ++a) acc += cget(i, k, a) * y[k][a];
        return acc;
      };
      auto gdot = [&](int i) -> double {
        reload_b();
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < CPW; ++k)
#pragma unroll
          for (int a = 0; a < 3; ++a) acc += cget(i, k, a) * (double)b[k][a];
        return acc;
      };
      double rrp = 0.0;  // |g - H c_0|^2 (plain start)
#pragma unroll
      for (int k = 0; k < CPW; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) rrp += r[k][a] * r[k][a];
      // Gram entries (upper triangle) and right-hand sides, reduced in
      // groups of four between the SpMVs and kept in LDS (uniform)
      __shared__ double gal[28];
      auto keep4 = [&](double (&v)[4], int o0, int o1, int o2, int o3) {
        if (threadIdx.x == 0) {
          gal[o0] = v[0];
          gal[o1] = v[1];
          gal[o2] = v[2];
          gal[o3] = v[3];
        }
      };
      // gal index of G_ij (i <= j) and of f_i
      auto gi = [](int i, int j) { return i * 6 - i * (i - 1) / 2 + (j - i); };
      constexpr int F0 = 21;
      // without a learned guess: H c_2 .. H c_5, kept for the start's residual
      double hc[4][CPW][3];
      if (!S.guess) {
        // No learned guess (c_1 = 0: its Gram row and right-hand side are
        // zero and the Cholesky drops it).  Nothing is published (the
        // histories were written by the previous launch), so the reductions
        // carry values only and each candidate's SpMV runs inside the round
        // trip of the reduction before it.
        double y[CPW][3];  // H c_0 = g - r
        reload_b();
#pragma unroll
        for (int k = 0; k < CPW; ++k)
#pragma unroll
          for (int a = 0; a < 3; ++a) y[k][a] = live[k] ? (double)b[k][a] - r[k][a] : 0.0;
        const double g05 = cdot(5, y);
        const double f0 = gdot(0), f2 = gdot(2), f3 = gdot(3), f4 = gdot(4), f5 = gdot(5);
        double R1[4] = {cdot(0, y), cdot(2, y), cdot(3, y), cdot(4, y)};
        grid_post(gs, R1, false);
        spmv_check64(S.dx, hc[0]);  // H c_2
        const double g22 = cdot(2, hc[0]), g23 = cdot(3, hc[0]), g24 = cdot(4, hc[0]), g25 = cdot(5, hc[0]);
        if (!grid_wait(gs, R1)) return;
        keep4(R1, gi(0, 0), gi(0, 2), gi(0, 3), gi(0, 4));
        double R2[4] = {g05, f0, f2, f3};
        grid_post(gs, R2, false);
        spmv_check64(h0, hc[1]);  // H c_3
        const double g33 = cdot(3, hc[1]), g34 = cdot(4, hc[1]), g35 = cdot(5, hc[1]);
        if (!grid_wait(gs, R2)) return;
        keep4(R2, gi(0, 5), F0 + 0, F0 + 2, F0 + 3);
        double R3[4] = {f4, f5, g22, g23};
        grid_post(gs, R3, false);
        spmv_check64(h1, hc[2]);  // H c_4
        const double g44 = cdot(4, hc[2]), g45 = cdot(5, hc[2]);
        if (!grid_wait(gs, R3)) return;
        keep4(R3, F0 + 4, F0 + 5, gi(2, 2), gi(2, 3));
        double R4[4] = {g24, g25, g33, g34};
        grid_post(gs, R4, false);
        double g55 = 0.0;
#pragma unroll
        for (int k = 0; k < CPW; ++k)
#pragma unroll
          for (int a = 0; a < 3; ++a) hc[3][k][a] = 0.0;
        if (has5) {
          spmv_check64(h5, hc[3]);  // H c_5
          g55 = cdot(5, hc[3]);
        }
        if (!grid_wait(gs, R4)) return;
        keep4(R4, gi(2, 4), gi(2, 5), gi(3, 3), gi(3, 4));
        double R5[4] = {g35, g44, g45, g55};
        if (!grid_reduce(gs, R5)) return;
        keep4(R5, gi(3, 5), gi(4, 4), gi(4, 5), gi(5, 5));
        if (threadIdx.x == 0) {
          gal[gi(0, 1)] = 0.0;
#pragma unroll
          for (int j = 1; j < 6; ++j) gal[gi(1, j)] = 0.0;
          gal[F0 + 1] = 0.0;
        }
      } else {
        {
          double* gb = pxbuf(pubx);
#pragma unroll
          for (int k = 0; k < CPW; ++k) {
            if (!own[k]) continue;
            const int row = chunk[k] * 64 + lane;
#pragma unroll
            for (int a = 0; a < 3; ++a) pub_store(&gb[3 * row + a], tagx(cget(1, k, a), pubx));
          }
          ++pubx;
        }
        {
          double y[CPW][3];  // H c_0 = g - r
          reload_b();
#pragma unroll
          for (int k = 0; k < CPW; ++k)
#pragma unroll
            for (int a = 0; a < 3; ++a) y[k][a] = live[k] ? (double)b[k][a] - r[k][a] : 0.0;
          double P[4] = {cdot(0, y), cdot(1, y), cdot(2, y), cdot(3, y)};
          double Q[4] = {cdot(4, y), cdot(5, y), gdot(0), gdot(1)};
          double Rr[4] = {gdot(2), gdot(3), gdot(4), gdot(5)};
          if (!grid_reduce(gs, P)) return;  // (also publishes the learned guess)
          keep4(P, gi(0, 0), gi(0, 1), gi(0, 2), gi(0, 3));
          if (!grid_reduce(gs, Q)) return;
          keep4(Q, gi(0, 4), gi(0, 5), F0 + 0, F0 + 1);
          if (!grid_reduce(gs, Rr)) return;
          keep4(Rr, F0 + 2, F0 + 3, F0 + 4, F0 + 5);
        }
        {
          // (each grid reduction's workgroup barrier also frees the staged
          // SpMV's line buffer for the next SpMV)
          double y[CPW][3];
          spmv_check64(pxbuf(pubx - 1), y);  // H c_1
          double A4[4] = {cdot(1, y), cdot(2, y), cdot(3, y), cdot(4, y)};
          double g15 = cdot(5, y);
          if (!grid_reduce(gs, A4)) return;
          keep4(A4, gi(1, 1), gi(1, 2), gi(1, 3), gi(1, 4));
          spmv_check64(S.dx, y);  // H c_2
          double B4[4] = {g15, cdot(2, y), cdot(3, y), cdot(4, y)};
          double g25 = cdot(5, y);
          if (!grid_reduce(gs, B4)) return;
          keep4(B4, gi(1, 5), gi(2, 2), gi(2, 3), gi(2, 4));
          spmv_check64(h0, y);  // H c_3
          double C4[4] = {g25, cdot(3, y), cdot(4, y), cdot(5, y)};
          if (!grid_reduce(gs, C4)) return;
          keep4(C4, gi(2, 5), gi(3, 3), gi(3, 4), gi(3, 5));
          spmv_check64(h1, y);  // H c_4
          double g44 = cdot(4, y), g45 = cdot(5, y);
          __syncthreads();
          double g55 = 0.0;
          if (has5) {
            spmv_check64(h5, y);  // H c_5
            g55 = cdot(5, y);
          }
          double D4[4] = {g44, g45, g55, 0.0};
          if (!grid_reduce(gs, D4)) return;
          keep4(D4, gi(4, 4), gi(4, 5), gi(5, 5), 27);
        }
      }
      __syncthreads();
      double Gm[6][6], fv[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        fv[i] = gal[F0 + i];
#pragma unroll
        for (int j = i; j < 6; ++j) Gm[i][j] = gal[gi(i, j)];
      }
      // pivoted Cholesky of the Gram matrix (same scalars in every workgroup);
      // the factor L (lower triangle, diagonal included) overwrites Gm in
      // place: Gm holds the Gram entries in its upper triangle
      double z[6], cf[6];
      bool kept[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double gjj = Gm[j][j];
        double d = gjj;
#pragma unroll
        for (int q = 0; q < j; ++q) d -= kept[q] ? Gm[j][q] * Gm[j][q] : 0.0;
        kept[j] = gjj > 0.0 && d > 1e-10 * gjj;
        Gm[j][j] = kept[j] ? sqrt(d) : 1.0;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
          double v = Gm[j][i];
#pragma unroll
          for (int q = 0; q < j; ++q) v -= kept[q] ? Gm[i][q] * Gm[j][q] : 0.0;
          Gm[i][j] = kept[j] ? v / Gm[j][j] : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        double v = fv[j];
#pragma unroll
        for (int q = 0; q < j; ++q) v -= Gm[j][q] * z[q];
        z[j] = kept[j] ? v / Gm[j][j] : 0.0;
      }
#pragma unroll
      for (int j = 5; j >= 0; --j) {
        double v = z[j];
#pragma unroll
        for (int i = j + 1; i < 6; ++i) v -= Gm[i][j] * cf[i];
        cf[j] = kept[j] ? v / Gm[j][j] : 0.0;
      }
      if (!S.guess) {
        // x0 = C a and its residual from the candidates' products:
        // g - H x0 = g - a_0 (g - r) - sum_{i >= 2} a_i H c_i (no
        // publication, no SpMV); |g - H c_0|^2 is summed by the restart's
        // reduction
        reload_b();
#pragma unroll
        for (int k = 0; k < CPW; ++k)
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            x[k][a] = cf[0] * cget(0, k, a) + cf[2] * cget(2, k, a) + cf[3] * cget(3, k, a) +
                      cf[4] * cget(4, k, a) + cf[5] * cget(5, k, a);
            const double hx = cf[0] * ((double)b[k][a] - r[k][a]) + cf[2] * hc[0][k][a] + cf[3] * hc[1][k][a] +
                              cf[4] * hc[2][k][a] + cf[5] * hc[3][k][a];
            r[k][a] = live[k] ? (double)b[k][a] - hx : 0.0;
          }
        guess_used = 3;
        rs_extra = rrp;
      } else {
        // x0 = C a, published for its residual
        double* xb = pxbuf(pubx);
#pragma unroll
        for (int k = 0; k < CPW; ++k) {
          double xn[3];
#pragma unroll
          for (int a = 0; a < 3; ++a)
            xn[a] = cf[0] * cget(0, k, a) + cf[1] * cget(1, k, a) + cf[2] * cget(2, k, a) + cf[3] * cget(3, k, a) +
                    cf[4] * cget(4, k, a) + cf[5] * cget(5, k, a);
#pragma unroll
          for (int a = 0; a < 3; ++a) x[k][a] = xn[a];
          if (own[k]) {
            const int row = chunk[k] * 64 + lane;
#pragma unroll
            for (int a = 0; a < 3; ++a) pub_store(&xb[3 * row + a], tagx(x[k][a], pubx));
          }
        }
        ++pubx;
        double red0[1] = {rrp};
        if (!grid_reduce(gs, red0)) return;
        rrp = red0[0];
        {
          double y[CPW][3];
          spmv_check64(xb, y);
          reload_b();
#pragma unroll
          for (int k = 0; k < CPW; ++k)
#pragma unroll
            for (int a = 0; a < 3; ++a) r[k][a] = live[k] ? (double)b[k][a] - y[k][a] : 0.0;
          guess_used = 3;
          guess_rrp = rrp;
        }
      }
    }
    if (!restart(rz, rr)) return;
    if (guess_used == 3 && !S.guess) guess_rrp = rs_extra;
    rs_extra = 0.0;
    // (rr: |g - H x0|^2 of the start, reduced by restart)
    if (guess_used == 3) guess_ratio = guess_rrp > 0.0 ? sqrt(rr / guess_rrp) : 0.0;
    bool first = true;
    double beta = 0.0;
    // pipelined: m = M w already published by the previous update, with the
    // dot products of the new vectors (the update issues the stores first
    // and overlaps the rest of its work with their completion)
    bool pre = false;
    double gamma_pre = 0.0, delta_pre = 0.0, rr_pre = 0.0;
    // pipelined: 1 / gamma and gamma * alpha of the previous iteration, one
    // copy per wave in LDS (wave-uniform; held in registers they are spilled
    // to scratch, whose reload would sit on the critical path after each
    // reduction).  With them
    //   beta = gamma / gamma_prev = gamma * (1 / gamma_prev),
    //   alpha = gamma / (delta - beta gamma / alpha_prev)
    //         = gamma P / (delta P - gamma^2),  P = gamma_prev alpha_prev,
    // one division after the reduction instead of three in sequence.
    __shared__ double carried[kWaves][2];
    const int cw = threadIdx.x >> 6;
    PCG_LAP(5);
    for (;;) {
      double gamma = 0.0, delta = 0.0;
      double n[CPW][3];  // pipelined: H m
      if constexpr (PIPE) {
        // gamma = r.u, delta = w.u, rr = r.r; m = M w published with them
        rr = 0.0;
        double* mb = pxbuf(pubx);
        float* mb32 = pzbuf(pub);
        if (pre) {
          gamma = gamma_pre;
          delta = delta_pre;
          rr = rr_pre;
        }
#pragma unroll
        for (int k = 0; k < CPW; ++k) {
          if (!own[k] || pre) continue;  // (uniform over the wave)
          const int row = chunk[k] * 64 + lane;
          double m[3];
          if constexpr (COARSE) apply_Mc(w[k], ccw, m);
          else apply_M<CPW, double>(M, k, w[k], m);
          if constexpr (PUB32) {
#pragma unroll
            for (int a = 0; a < 3; ++a) pub_store(&mb32[3 * row + a], (float)m[a]);
          } else {
            // the chunk's 192 values go out as three contiguous 512-byte
            // stores (transposed through this wave's LDS slice; LDS
            // operations of a wave complete in order) instead of three
            // 24-byte-strided ones that each touch 24 lines partially
#pragma unroll
            for (int a = 0; a < 3; ++a) trw[3 * lane + a] = m[a];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int j = 0; j < 3; ++j) pub_store(&mb[(size_t)chunk[k] * 192 + lane + 64 * j], tagx(trw[lane + 64 * j], pubx));
            __builtin_amdgcn_wave_barrier();
          }
          // u = M r: held by recurrence in the FP32-published form, else
          // recomputed from r where used (exact in exact arithmetic; fewer
          // registers and no drift between u and M r)
          double ul[3];
          if constexpr (PUB32) {
#pragma unroll
            for (int a = 0; a < 3; ++a) ul[a] = u[k][a];
          } else if constexpr (COARSE) {
#pragma unroll
            for (int a = 0; a < 3; ++a) ul[a] = u[k][a];  // (M r, carried)
          } else {
            apply_M<CPW, double>(M, k, r[k], ul);
          }
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            gamma += r[k][a] * ul[a];
            delta += w[k][a] * ul[a];
            rr += r[k][a] * r[k][a];
          }
        }
        if constexpr (PUB32) ++pub;
        else if (!pre) ++pubx;
        pre = false;
        PCG_LAP(2);
        double red[3] = {gamma, delta, rr};
        if (ovl) {
          // n = H m runs inside the reduction's round trip
          grid_post(gs, red, false);
          if (!spmv_tagged(pxbuf(pubx - 1), (pubx - 1) & 3, n)) return;
          PCG_LAP(0);
          if (!grid_wait(gs, red)) return;
        } else if (!grid_reduce(gs, red)) return;
        gamma = red[0];
        delta = red[1];
        rr = red[2];
        PCG_LAP(1);
      }
      if (rr <= tol2_rec || it >= S.cg_maxiter) {
        // true residual with the FP64 iterate
        double* xb = pxbuf(pubx);
#pragma unroll
        for (int k = 0; k < CPW; ++k) {
          if (!own[k]) continue;
          const int row = chunk[k] * 64 + lane;
#pragma unroll
          for (int a = 0; a < 3; ++a) pub_store(&xb[3 * row + a], tagx(x[k][a], pubx));
        }
        ++pubx;
        double red0[1] = {0.0};
        if (!grid_reduce(gs, red0)) return;
        double acc = 0.0;
        double rt[CPW][3];
        {
          double y[CPW][3];
          spmv_check64(xb, y);
          reload_b();
#pragma unroll
          for (int k = 0; k < CPW; ++k)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              rt[k][a] = live[k] ? (double)b[k][a] - y[k][a] : 0.0;
              acc += rt[k][a] * rt[k][a];
            }
        }
        double redt[1] = {acc};
        if (!grid_reduce(gs, redt)) return;
        rt2 = redt[0];
        PCG_LAP(4);
        if (rt2 <= tol2 || it >= S.cg_maxiter || restarts >= 16) break;
        ++restarts;
#pragma unroll
        for (int k = 0; k < CPW; ++k)
#pragma unroll
          for (int a = 0; a < 3; ++a) r[k][a] = rt[k][a];
        if (!restart(rz, rr)) return;
        // restart the Krylov sequence from the replaced residual (keeping the
        // old direction diverges when the residual gap is large)
        first = true;
        beta = 0.0;
        PCG_LAP(4);
        continue;
      }
      if constexpr (PIPE) {
        // n = H m (unless overlapped above); scalars; all vector recurrences
        // local (FP64)
        if constexpr (PUB32) {
          float y[CPW][3];
          spmv_rows<CPW, RW>(S, chunk, own, lane, rb, pzbuf(pub - 1), y);
#pragma unroll
          for (int k = 0; k < CPW; ++k)
#pragma unroll
            for (int a = 0; a < 3; ++a) n[k][a] = (double)y[k][a];
        } else if (!ovl) {
          spmv_resident64(pxbuf(pubx - 1), n);
        }
        if (!ovl) PCG_LAP(0);
        double alpha;
        if (first) {
          beta = 0.0;
          alpha = gamma / delta;
        } else {
          const double P = carried[cw][1];
          beta = gamma * carried[cw][0];
          alpha = gamma * P / (delta * P - gamma * gamma);
        }
        if (!(alpha > 0.0) || !isfinite(alpha)) {  // breakdown: fall to the residual check
          it = S.cg_maxiter;
          continue;
        }
        // the next iteration's carried scalars, computed here so that the
        // division overlaps the vector updates
        const double ig_next = 1.0 / gamma, P_next = gamma * alpha;
        if constexpr (COARSE) {
          // as below, with the coarse coefficients: those of the new w from
          // one workgroup sum, those of s and r by the same recurrences as
          // the vectors (linear: c(s') = c(w) + beta c(s), c(r') = c(r) -
          // alpha c(s')).  The new w's share of the sum goes out first, the
          // other recurrences and dot products run while the other waves
          // arrive, then m = M w is published.  Every wave takes part.
          double* mb = pxbuf(pubx);
          double g2 = 0.0, d2 = 0.0, r2 = 0.0;
          double wo[3], mn[3], ul[3], cwn[6], csn[6], crn[6];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            wo[a] = w[0][a];
            zd[0][a] = n[0][a] + beta * zd[0][a];
            w[0][a] -= alpha * zd[0][a];
          }
          coarse_post(w[0]);
          // u = M r of the current r is carried (set by the restart and by
          // the previous update): one application of M fewer per iteration
#pragma unroll
          for (int j = 0; j < 6; ++j) {
            csn[j] = ccw[j] + beta * ccs[j];
            crn[j] = ccr[j] - alpha * csn[j];
          }
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            sd[0][a] = wo[a] + beta * sd[0][a];
            pd[0][a] = u[0][a] + beta * pd[0][a];
            x[0][a] += alpha * pd[0][a];
            r[0][a] -= alpha * sd[0][a];
          }
          apply_Mc(r[0], crn, ul);
#pragma unroll
          for (int a = 0; a < 3; ++a) u[0][a] = ul[a];
          if (own[0]) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              g2 += r[0][a] * ul[a];
              d2 += w[0][a] * ul[a];
              r2 += r[0][a] * r[0][a];
            }
          }
          coarse_get(cwn);
          if (own[0]) {
            apply_Mc(w[0], cwn, mn);
#pragma unroll
            for (int a = 0; a < 3; ++a) trw[3 * lane + a] = mn[a];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int j = 0; j < 3; ++j) pub_store(&mb[(size_t)chunk[0] * 192 + lane + 64 * j], tagx(trw[lane + 64 * j], pubx));
            __builtin_amdgcn_wave_barrier();
          }
#pragma unroll
          for (int j = 0; j < 6; ++j) {
            ccw[j] = cwn[j];
            ccs[j] = csn[j];
            ccr[j] = crn[j];
          }
          ++pubx;
          pre = true;
          gamma_pre = g2;
          delta_pre = d2;
          rr_pre = r2;
        } else if constexpr (!PUB32) {
          // w first; the next m = M w goes out at once, and the remaining
          // recurrences and the next dot products run while it drains
          double* mb = pxbuf(pubx);
          double g2 = 0.0, d2 = 0.0, r2 = 0.0;
#pragma unroll
          for (int k = 0; k < CPW; ++k) {
            if (!own[k]) continue;  // (uniform over the wave)
            double wo[3], mn[3], ul[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              wo[a] = w[k][a];
              zd[k][a] = n[k][a] + beta * zd[k][a];
              w[k][a] -= alpha * zd[k][a];
            }
            apply_M<CPW, double>(M, k, w[k], mn);
#pragma unroll
            for (int a = 0; a < 3; ++a) trw[3 * lane + a] = mn[a];
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int j = 0; j < 3; ++j) pub_store(&mb[(size_t)chunk[k] * 192 + lane + 64 * j], tagx(trw[lane + 64 * j], pubx));
            __builtin_amdgcn_wave_barrier();
            // u = M r before and after the r update (no q recurrence)
            apply_M<CPW, double>(M, k, r[k], ul);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              sd[k][a] = wo[a] + beta * sd[k][a];
              pd[k][a] = ul[a] + beta * pd[k][a];
              x[k][a] += alpha * pd[k][a];
              r[k][a] -= alpha * sd[k][a];
            }
            apply_M<CPW, double>(M, k, r[k], ul);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              g2 += r[k][a] * ul[a];
              d2 += w[k][a] * ul[a];
              r2 += r[k][a] * r[k][a];
            }
          }
          ++pubx;
          pre = true;
          gamma_pre = g2;
          delta_pre = d2;
          rr_pre = r2;
        } else {
#pragma unroll
          for (int k = 0; k < CPW; ++k) {
            if (!own[k]) continue;
            double m[3];
            apply_M<CPW, double>(M, k, w[k], m);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              zd[k][a] = n[k][a] + beta * zd[k][a];
              qd[k][a] = m[a] + beta * qd[k][a];
              sd[k][a] = w[k][a] + beta * sd[k][a];
              pd[k][a] = u[k][a] + beta * pd[k][a];
              x[k][a] += alpha * pd[k][a];
              r[k][a] -= alpha * sd[k][a];
              u[k][a] -= alpha * qd[k][a];
              w[k][a] -= alpha * zd[k][a];
            }
          }
        }
        PCG_LAP(2);
        if (lane == 0) {
          carried[cw][0] = ig_next;
          carried[cw][1] = P_next;
        }
        first = false;
        ++it;
        continue;
      } else {
        // phase 1: q = H p via y = H z
        double pq = 0.0;
        const float fb = (float)beta;
        {
          float y[CPW][3];
          spmv_rows<CPW, RW, LOWREG ? 4 : 8>(S, chunk, own, lane, rb, pzbuf(pub - 1), y);
#pragma unroll
          for (int k = 0; k < CPW; ++k) {
            if (!own[k]) continue;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              q[k][a] = y[k][a] + fb * q[k][a];
              p[k][a] = z[k][a] + fb * p[k][a];
              pq += (double)p[k][a] * q[k][a];
            }
          }
        }
        PCG_LAP(0);
        double redp[1] = {pq};
        if (!grid_reduce(gs, redp)) return;
        pq = redp[0];
        PCG_LAP(1);
        if (!(pq > 0.0)) {  // breakdown (non-SPD or exact solution): fall to the residual check
          it = S.cg_maxiter;
          continue;
        }
        const double alpha = rz / pq;
        double rz2 = 0.0;
        rr = 0.0;
        float* zb = pzbuf(pub);
#pragma unroll
        for (int k = 0; k < CPW; ++k) {
          if (!own[k]) continue;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            x[k][a] += alpha * (double)p[k][a];
            r[k][a] -= alpha * (double)q[k][a];
          }
        }
        precondition();
#pragma unroll
        for (int k = 0; k < CPW; ++k) {
          if (!own[k]) continue;
          const int row = chunk[k] * 64 + lane;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            pub_store(&zb[3 * row + a], z[k][a]);
            rz2 += r[k][a] * z[k][a];
            rr += r[k][a] * r[k][a];
          }
        }
        ++pub;
        PCG_LAP(2);
        double redz[2] = {rz2, rr};
        if (!grid_reduce(gs, redz)) return;
        rz2 = redz[0];
        rr = redz[1];
        PCG_LAP(3);
        beta = rz2 / rz;
        rz = rz2;
        ++it;
      }
    }
  }
  // write the solution (g == 0 gives dx = 0); the previous two steps move
  // down the history (no other workgroup reads these rows any more: every
  // workgroup has passed the final reduction)
#pragma unroll
  for (int k = 0; k < CPW; ++k) {
    if (!live[k]) continue;
    const int row = chunk[k] * 64 + lane;
    if (S.dx_hist) {
      double* h0 = S.dx_hist;
      double* h1 = S.dx_hist + nrow_pad * 3;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        h1[3 * row + a] = h0[3 * row + a];
        h0[3 * row + a] = S.dx[3 * row + a];
      }
    }
    if (S.dx_same && newton_iter >= 0 && newton_iter < S.dx_same_iters) {
      double* hs = S.dx_same + (size_t)newton_iter * nrow_pad * 3;
#pragma unroll
      for (int a = 0; a < 3; ++a) hs[3 * row + a] = bnorm2 > 0.0 ? x[k][a] : 0.0;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) S.dx[3 * row + a] = bnorm2 > 0.0 ? x[k][a] : 0.0;
  }
  if (blockIdx.x == wg0 && threadIdx.x == 0) {
    if (newton_iter >= 0 && newton_iter < MFEM_MAX_STATS_ITERS) {
      double* st = S.stats + ((size_t)body * MFEM_MAX_STATS_ITERS + newton_iter) * MFEM_STATS_PER_ITER;
      st[0] = (double)it;
      st[1] = bnorm2 > 0.0 ? sqrt(rt2 / bnorm2) : 0.0;
      st[7] = (double)restarts;
#pragma unroll
      for (int k = 0; k < 6; ++k) st[8 + k] = (double)tk[k];
      st[14] = (double)guess_used;
      st[15] = guess_ratio;
    }
    *ep_ptr = (int32_t)gs.epoch;
  }
#undef PCG_LAP
}

template <int CPW, int RW, int MODE>
__global__ void __launch_bounds__(kBlock) k_pcg(mfem_solver S, int newton_iter) {
  pcg_body<CPW, RW, MODE, false>(S, newton_iter);
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_pcg_lowreg(mfem_solver S, int newton_iter) {
  pcg_body<1, 0, 0, true>(S, newton_iter);
}

// ---------------------------------------------------------------------------
// learned warm start: candidate guesses and per-vertex features
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__global__ void __launch_bounds__(kBlock) k_ws_features(mfem_solver S, float dt, int newton_iter) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.n_verts) return;
  const int c = i >> 6, lane = i & 63;
  float c1[3], r1[3], ring = 0.f, D[9];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    c1[a] = S.x_tilde[3 * i + a] - S.x[3 * i + a];
    r1[a] = S.g[3 * i + a];
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) D[q] = S.sell_val[val_index(S.diag_lane[i], q)];
  int valence = 0;
  for (int sl = S.chunk_ptr[c]; sl < S.chunk_ptr[c + 1]; ++sl) {
    const int j = S.sell_col[sl * 64 + lane];
    const float* vb = S.sell_val + (size_t)sl * 576 + lane;
    float bv[9], fro = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      bv[q] = vb[64 * q];
      fro += bv[q] * bv[q];
    }
    float cj[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) cj[a] = S.x_tilde[3 * j + a] - S.x[3 * j + a];
#pragma unroll
    for (int a = 0; a < 3; ++a) r1[a] -= bv[3 * a] * cj[0] + bv[3 * a + 1] * cj[1] + bv[3 * a + 2] * cj[2];
    if (fro > 0.f && j != i) {
      ring += sqrtf(fro);
      ++valence;
    }
  }
  float Mi[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) Mi[q] = (q % 4 == 0) ? 1.f : 0.f;
  if (S.precond && fabsf(det3(D)) >= S.sing_thr) inv3(D, Mi);
  float c2[3], c3[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    c2[a] = Mi[3 * a] * r1[0] + Mi[3 * a + 1] * r1[1] + Mi[3 * a + 2] * r1[2];
    c3[a] = (float)S.dx[3 * i + a];
  }
  auto nrm = [](const float* v) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
  auto cosv = [&](const float* u, const float* v) {
    const float d = nrm(u) * nrm(v);
    return d > 0.f ? (u[0] * v[0] + u[1] * v[1] + u[2] * v[2]) / d : 0.f;
  };
  const float n1 = nrm(c1), n2 = nrm(c2), n3 = nrm(c3);
  const float sc = n1 + n2 + n3 + 1e-30f;
  float dfro = 0.f, hfro = 0.f;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    dfro += D[q] * D[q];
    hfro += S.Hv[9 * i + q] * S.Hv[9 * i + q];
  }
  dfro = sqrtf(dfro) + 1e-30f;
  hfro = sqrtf(hfro);
  const float im = S.inv_mass[i];
  const float mh = (im == 0.f ? 1.f + S.attach_k : 1.f / im) / (dt * dt);
  float g[3] = {S.g[3 * i], S.g[3 * i + 1], S.g[3 * i + 2]};
  const float ng = nrm(g);
  float f[16];
  f[0] = n1 / sc;
  f[1] = n2 / sc;
  f[2] = n3 / sc;
  f[3] = cosv(c1, c2);
  f[4] = cosv(c1, c3);
  f[5] = cosv(c2, c3);
  f[6] = hfro / dfro;
  f[7] = fminf(mh / (dfro / 1.7320508f), 10.f);
  f[8] = im == 0.f ? 1.f : 0.f;
  f[9] = (float)newton_iter / (float)(S.iterations > 0 ? S.iterations : 1);
  f[10] = ring / dfro;
  f[11] = fminf(ng / (dfro * sc), 10.f);
  f[12] = fminf(nrm(r1) / (ng + 1e-30f), 10.f);
  f[13] = (float)valence / 16.f;
  f[14] = 1.f;
  // step length the line search accepted in the previous Newton iteration
  const int body = S.n_bodies > 1 ? body_of(S.body_vert, S.n_bodies, i) : 0;
  f[15] = (newton_iter > 0 && newton_iter <= MFEM_MAX_STATS_ITERS)
              ? (float)S.stats[((size_t)body * MFEM_MAX_STATS_ITERS + newton_iter - 1) * MFEM_STATS_PER_ITER + 4]
              : 0.f;
#pragma unroll
  for (int k = 0; k < 16; ++k) S.ws_x[16 * i + k] = f2bf16(f[k]);
  const float is = 1.f / sc;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    S.ws_aux[10 * i + a] = c1[a] * is;
    S.ws_aux[10 * i + 3 + a] = c2[a] * is;
    S.ws_aux[10 * i + 6 + a] = c3[a] * is;
  }
  S.ws_aux[10 * i + 9] = sc;
}

__global__ void __launch_bounds__(kBlock) k_ws_apply(mfem_solver S) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.n_verts) return;
  const float* a = S.ws_aux + 10 * i;
  const float* y = S.ws_y + 3 * i;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    S.guess[3 * i + d] = a[9] * ((1.f + y[2]) * a[6 + d] + y[0] * a[d] + y[1] * a[3 + d]);
}

// ---------------------------------------------------------------------------
// features and weights of the coarse-weight network (one workgroup per
// aggregate: rows [256 b, 256 b + 256), the PCG workgroup's rows)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_coarse_features(mfem_solver S, float dt) {
  __shared__ double red[2][kWaves][24];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool lv = i < S.n_verts;
  float xi[3] = {0.f, 0.f, 0.f};
  if (lv) load3(S.x, i, xi);
  double m4[4] = {xi[0], xi[1], xi[2], lv ? 1.0 : 0.0};
  wg_sum<4>(m4, red[0]);
  const double cnt = m4[3];
  double d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) d[a] = (lv && cnt > 0.0) ? (double)xi[a] - m4[a] / cnt : 0.0;
  double r2[1] = {d[0] * d[0] + d[1] * d[1] + d[2] * d[2]};
  wg_sum<1>(r2, red[1]);
  if (!lv) return;  // (the network runs over the active vertices only)
  const double rad = (cnt > 0.0 && r2[0] > 0.0) ? sqrt(r2[0] / cnt) : 1.0;
  const int dl = S.diag_lane[i];
  float D[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) D[q] = S.sell_val[val_index(dl, q)];
  const float trD = D[0] + D[4] + D[8];
  float nD = 0.f;
#pragma unroll
  for (int q = 0; q < 9; ++q) nD += D[q] * D[q];
  nD = sqrtf(nD);
  const float im = S.inv_mass[i];
  const float md = (im == 0.f ? 1.f + S.attach_k : 1.f / im) / (dt * dt);
  const float trHv = S.Hv[9 * i] + S.Hv[9 * i + 4] + S.Hv[9 * i + 8];
  const int c = i >> 6, lane = i & 63;
  float off_all = 0.f, off_out = 0.f;
  int nblk = 0, nout = 0;
  for (int sl = S.chunk_ptr[c]; sl < S.chunk_ptr[c + 1]; ++sl) {
    const int col = S.sell_col[sl * 64 + lane];
    const float* vb = S.sell_val + (size_t)sl * 576 + lane;
    float nb = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) nb += vb[64 * q] * vb[64 * q];
    if (!(nb > 0.f) || col == i) continue;
    nb = sqrtf(nb);
    const bool out = (col >> 8) != (i >> 8);
    off_all += nb;
    ++nblk;
    if (out) {
      off_out += nb;
      ++nout;
    }
  }
  const float tiny = 1e-30f;
  float f[16];
  f[0] = 1.f;
  f[1] = off_out / fmaxf(off_all, tiny);
  f[2] = 3.f * md / fmaxf(trD, tiny);
  f[3] = trHv / fmaxf(trD, tiny);
  f[4] = (float)(sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / rad);
  f[5] = off_all / fmaxf(nD, tiny);
  f[6] = im == 0.f ? 1.f : 0.f;
  f[7] = 0.25f * log10f(fmaxf(trD, tiny) / fmaxf(3.f * md, tiny));
  f[8] = (float)(nblk + 1) / 16.f;
  f[9] = (float)nout / 16.f;
  f[10] = (S.n_shapes > 0 && S.dist[i] < S.d1) ? 1.f : 0.f;
#pragma unroll
  for (int k = 11; k < 16; ++k) f[k] = 0.f;
#pragma unroll
  for (int k = 0; k < 16; ++k) S.cw_x[16 * (size_t)i + k] = f2bf16(f[k]);
}

__global__ void __launch_bounds__(kBlock) k_coarse_weights(mfem_solver S) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.n_verts) return;
  S.coarse_w[i] = fminf(fmaxf(S.cw_y[3 * i], 0.f), 2.f);
}

// ---------------------------------------------------------------------------
// local solve (kernels.py:252-283)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_local(mfem_solver S) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= S.n_tets) return;
  int4 tv = reinterpret_cast<const int4*>(S.tets)[e];
  const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
  float d[12];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int k = 0; k < 3; ++k) d[3 * a + k] = (float)S.dx[3 * vid[a] + k];
  float t[6];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    float acc = S.gl[6 * e + r];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc -= S.G[72 * e + r * 12 + k] * d[k];
    t[r] = acc;
  }
  float lam[6], w[6];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc += S.Hl[21 * e + sym21(r, c)] * t[c];
    lam[r] = -acc;
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) w[c] = SYMD[c] * lam[c] + S.gs[6 * e + c];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    float acc = 0.f;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc += S.Hsinv[21 * e + sym21(r, c)] * w[c];
    S.ds[6 * e + r] = -acc;
    S.lam[6 * e + r] = lam[r];
  }
}

// ---------------------------------------------------------------------------
// line search (line_search.py:76-163; solver.py:803-928)
// ---------------------------------------------------------------------------
// kinetic / attachment / barrier / friction terms of vertex i at x + a dx
__device__ __forceinline__ double vertex_merit(const mfem_solver& S, int i, float a, float dt) {
  float x[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) x[k] = S.x[3 * i + k] + a * (float)S.dx[3 * i + k];
  const double idt2 = 1.0 / ((double)dt * (double)dt);
  double d2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double d = (double)(x[k] - S.x_tilde[3 * i + k]);
    d2 += d * d;
  }
  double f = 0.5 * (double)S.mass[i] * d2 * idt2;
  if (S.inv_mass[i] == 0.f) {
    double p2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double d = (double)(x[k] - S.pin_pos[3 * i + k]);
      p2 += d * d;
    }
    f += 0.5 * (double)S.attach_k * p2 * idt2;
  }
  if (S.n_shapes > 0) {
    float grad[3];
    int sid;
    double d;
    shape_distance(S, x, grad, &sid, &d);
    if (d < S.d1) {
      double e, de, d2e;
      barrier64(d, S.d0, S.d1, e, de, d2e);
      f += (double)S.kappa * e;
    }
    f += (double)friction_eval(S, i, x, dt, nullptr, nullptr);
  }
  return f;
}

// Psi(s + a ds) - lambda . c(x + a dx, s + a ds)
__device__ __forceinline__ double tet_merit(const mfem_solver& S, int e, float a) {
  int4 tv = reinterpret_cast<const int4*>(S.tets)[e];
  const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
  float xv[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int k = 0; k < 3; ++k) xv[v][k] = S.x[3 * vid[v] + k] + a * (float)S.dx[3 * vid[v] + k];
  float s[6], D[9];
#pragma unroll
  for (int k = 0; k < 6; ++k) s[k] = S.s[6 * e + k] + a * S.ds[6 * e + k];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = S.dminv[9 * e + k];
  const float mu = S.materials[3 * e], la = S.materials[3 * e + 1], vol = S.vol[e];
  float psi = (S.tet_energy ? S.tet_energy[e] : S.energy) == 0 ? arap_energy(s, mu, vol) : nh_energy(s, mu, la, vol);
  float F[9], U[9], sg[3], V[9], S6[6];
  deformation_gradient(xv[0], xv[1], xv[2], xv[3], D, F);
  svd3(F, U, sg, V);
  stretch_sym6(sg, V, S6);
  double lc = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) lc += (double)(SYMD[k] * (S6[k] - s[k])) * (double)S.lam[6 * e + k];
  return (double)psi - lc;
}

// triangle-capsule barrier energy at x + a dx, weighted by the barycentric
// weights of the last full evaluation (contact.py:1050-1088)
__device__ __forceinline__ double tri_merit(const mfem_solver& S, int t, float a) {
  float v[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int i = S.tris[3 * t + c];
#pragma unroll
    for (int k = 0; k < 3; ++k) v[c][k] = S.x[3 * i + k] + a * (float)S.dx[3 * i + k];
  }
  float w[3], n[3];
  int sid;
  double d;
  tri_capsule(S, v[0], v[1], v[2], w, n, &sid, &d);
  if (!(d < S.d1)) return 0.0;
  double e, de, d2e;
  barrier64(d, S.d0, S.d1, e, de, d2e);
  const float* wl = S.tri_bary + 3 * t;
  return (double)S.kappa * e * ((double)wl[0] + (double)wl[1] + (double)wl[2]);
}

// (g_x + grad B + grad D) . dx for vertex i: the vertex part of the slope
__device__ __forceinline__ double vertex_slope_own(const mfem_solver& S, int i) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    acc += (double)(S.gx[3 * i + k] + S.gB[3 * i + k] + S.gD[3 * i + k]) * S.dx[3 * i + k];
  if (self_contact_on(S))
#pragma unroll
    for (int k = 0; k < 3; ++k) acc += S.sc_g[3 * i + k] * S.dx[3 * i + k];
  if (shell_on(S))
#pragma unroll
    for (int k = 0; k < 3; ++k) acc += (double)S.gS[3 * i + k] * S.dx[3 * i + k];
  return acc;
}

// -(G^T lambda)_e . dx over the tet's four vertices: the multiplier part of
// the vertices' slopes, summed per tet (each G record read once, whole)
__device__ __forceinline__ double tet_lambda_slope(const mfem_solver& S, int e) {
  const int4 tv = reinterpret_cast<const int4*>(S.tets)[e];
  const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
  float lam[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) lam[q] = S.lam[6 * e + q];
  double acc = 0.0;
#pragma unroll 1
  for (int a = 0; a < 4; ++a)  // (one vertex's 18 Jacobian entries at a time: registers)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      float t = 0.f;
#pragma unroll
      for (int q = 0; q < 6; ++q) t += S.G[72 * e + q * 12 + 3 * a + r] * lam[q];
      acc -= (double)t * S.dx[3 * vid[a] + r];
    }
  return acc;
}

__device__ __forceinline__ double tet_slope(const mfem_solver& S, int e) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k)
    acc += (double)(S.gs[6 * e + k] + SYMD[k] * S.lam[6 * e + k]) * (double)S.ds[6 * e + k];
  return acc;
}

// Fused local solve (fused_local): tet e's element at the current x (the
// records k_element<true> would store), its local solve
// lambda = -H_l (g_l - G dx_e), ds = -H_s^-1 (SYM lambda + g_s)
// (kernels.py:252-283; stored into lam and ds), and its line-search terms
// with the new multipliers: the merit at alpha = 0 and at a1, and the slope
// (g_s + SYM lambda) . ds - (G^T lambda) . dx_e -- each formed as tet_merit,
// tet_slope and tet_lambda_slope form them from the stored records.
__device__ __forceinline__ void tet_local_merit(const mfem_solver& S, int e, float a1, double& f0, double& sl, double& fa) {
  float xv[4][3];
  int4 tv;
  gather_tet(S, e, xv, tv);
  const int vid[4] = {tv.x, tv.y, tv.z, tv.w};
  float D[9], s[6];
#pragma unroll
  for (int q = 0; q < 9; ++q) D[q] = S.dminv[9 * e + q];
#pragma unroll
  for (int q = 0; q < 6; ++q) s[q] = S.s[6 * e + q];
  float lam[6], ds[6], S6[6], psi;
  double slope;
  {
    TetElement el;
    tet_element(S, e, xv, D, s, el);
    float d[12];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) d[3 * a + k] = (float)S.dx[3 * vid[a] + k];
    float t[6], w[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      float acc = el.gl[r];
#pragma unroll
      for (int k = 0; k < 12; ++k) acc -= el.G[r * 12 + k] * d[k];
      t[r] = acc;
    }
    // (the symmetric blocks are read from their upper triangle, as the
    // packed records are)
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      float acc = 0.f;
#pragma unroll
      for (int c = 0; c < 6; ++c) acc += el.Hl[r < c ? r * 6 + c : c * 6 + r] * t[c];
      lam[r] = -acc;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) w[c] = SYMD[c] * lam[c] + el.gs[c];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      float acc = 0.f;
#pragma unroll
      for (int c = 0; c < 6; ++c) acc += el.Hi[r < c ? r * 6 + c : c * 6 + r] * w[c];
      ds[r] = -acc;
    }
    slope = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) slope += (double)(el.gs[k] + SYMD[k] * lam[k]) * (double)ds[k];
    double ls = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        float tt = 0.f;
#pragma unroll
        for (int q = 0; q < 6; ++q) tt += el.G[q * 12 + 3 * a + r] * lam[q];
        ls -= (double)tt * S.dx[3 * vid[a] + r];
      }
    slope += ls;
#pragma unroll
    for (int k = 0; k < 6; ++k) S6[k] = el.S6[k];
    psi = el.psi;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    S.lam[6 * e + k] = lam[k];
    S.ds[6 * e + k] = ds[k];
  }
  double lc = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) lc += (double)(SYMD[k] * (S6[k] - s[k])) * (double)lam[k];
  f0 += (double)psi - lc;
  sl += slope;
  // merit at a1 (tet_merit with the new lambda and ds)
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int k = 0; k < 3; ++k) xv[v][k] = xv[v][k] + a1 * (float)S.dx[3 * vid[v] + k];
  float s1[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) s1[k] = s[k] + a1 * ds[k];
  const float mu = S.materials[3 * e], la = S.materials[3 * e + 1], vol = S.vol[e];
  const float psi1 = (S.tet_energy ? S.tet_energy[e] : S.energy) == 0 ? arap_energy(s1, mu, vol)
                                                                      : nh_energy(s1, mu, la, vol);
  float F[9], U[9], sg[3], V[9], S61[6];
  deformation_gradient(xv[0], xv[1], xv[2], xv[3], D, F);
  svd3(F, U, sg, V);
  stretch_sym6(sg, V, S61);
  double lc1 = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) lc1 += (double)(SYMD[k] * (S61[k] - s1[k])) * (double)lam[k];
  fa += (double)psi1 - lc1;
}

__global__ void __launch_bounds__(kBlock) k_linesearch(mfem_solver S, float dt, int newton_iter) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // see k_pcg
  // a batch: the body's workgroups search over the body's vertices, tets
  // and triangles only, with the body's own merit, step and stats
  const bool batch = S.n_bodies > 1;
  const int body = batch ? body_of(S.body_ls_wg, S.n_bodies, blockIdx.x) : 0;
  const int wg0 = batch ? S.body_ls_wg[body] : 0;
  const int wg1 = batch ? S.body_ls_wg[body + 1] : (int)gridDim.x;
  const int v0 = batch ? S.body_vert[body] : 0, v1 = batch ? S.body_vert[body + 1] : S.n_verts;
  const int e0 = batch ? S.body_tet[body] : 0, e1 = batch ? S.body_tet[body + 1] : S.n_tets;
  const int f0t = batch ? S.body_tri[body] : 0, f1t = batch ? S.body_tri[body + 1] : S.n_tris;
  int32_t* ep_ptr = batch ? S.body_epoch + body : S.status + MFEM_STATUS_EPOCH;
  GridSync gs{S.red, S.flags, S.status, (uint32_t)*ep_ptr, false, wg0, wg1, kRedRegionWords};
  const int tid = (blockIdx.x - wg0) * kBlock + threadIdx.x;
  const int nth = (wg1 - wg0) * kBlock;
  // the first pass also evaluates the first trial (alpha0): one pass and one
  // grid reduction fewer when it is accepted.  Each sum is formed in the same
  // order as in a trial pass of its own, so f1 is the same number.
  // with self-contact the first step is cut to the CCD bound (<= 1)
  const bool sc = self_contact_on(S);
  const double alpha0 = sc ? fmin((double)S.ls_alpha0, (double)S.sc_alpha[0]) : (double)S.ls_alpha0;
  const float a_first = (float)alpha0;
  double f = 0.0, sl = 0.0, fa = 0.0;
  for (int i = v0 + tid; i < v1; i += nth) {
    f += vertex_merit(S, i, 0.f, dt);
    sl += vertex_slope_own(S, i);
    fa += vertex_merit(S, i, a_first, dt);
  }
  if (S.fused_local) {
    // the local solve runs here, tet by tet (k_element<false> stored no records)
    for (int e = e0 + tid; e < e1; e += nth) tet_local_merit(S, e, a_first, f, sl, fa);
  } else {
    for (int e = e0 + tid; e < e1; e += nth) {
      f += tet_merit(S, e, 0.f);
      sl += tet_slope(S, e) + tet_lambda_slope(S, e);
      fa += tet_merit(S, e, a_first);
    }
  }
  if (tri_contact_on(S))
    for (int t = f0t + tid; t < f1t; t += nth) {
      f += tri_merit(S, t, 0.f);
      fa += tri_merit(S, t, a_first);
    }
  if (shell_on(S))
    for (int t = f0t + tid; t < f1t; t += nth) {
      f += shell_merit(S, t, 0.f);
      fa += shell_merit(S, t, a_first);
    }
  if (sc) {
    const int npt = sc_pair_count(S, 0), nee = sc_pair_count(S, 1);
    for (int k = tid; k < npt; k += nth) {
      f += sc_pair_merit<0>(S, k, 0.f);
      fa += sc_pair_merit<0>(S, k, a_first);
    }
    for (int k = tid; k < nee; k += nth) {
      f += sc_pair_merit<1>(S, k, 0.f);
      fa += sc_pair_merit<1>(S, k, a_first);
    }
  }
  double r0[3] = {f, sl, fa};
  if (!grid_reduce(gs, r0)) return;
  const double f0 = r0[0];
  const double t = -(double)S.ls_c * r0[1];
  double alpha = alpha0, used = alpha, f1 = f0;
  int it = 0;
  bool cond;
  do {
    if (it == 0) {
      f1 = r0[2];
    } else {
      const float a = (float)alpha;
      f = 0.0;
      for (int i = v0 + tid; i < v1; i += nth) f += vertex_merit(S, i, a, dt);
      for (int e = e0 + tid; e < e1; e += nth) f += tet_merit(S, e, a);
      if (tri_contact_on(S))
        for (int t = f0t + tid; t < f1t; t += nth) f += tri_merit(S, t, a);
      if (shell_on(S))
        for (int t = f0t + tid; t < f1t; t += nth) f += shell_merit(S, t, a);
      if (sc) {
        const int npt = sc_pair_count(S, 0), nee = sc_pair_count(S, 1);
        for (int k = tid; k < npt; k += nth) f += sc_pair_merit<0>(S, k, a);
        for (int k = tid; k < nee; k += nth) f += sc_pair_merit<1>(S, k, a);
      }
      double r1[1] = {f};
      if (!grid_reduce(gs, r1)) return;
      f1 = r1[0];
    }
    cond = ((f0 - f1) + (double)S.ls_thr < alpha * t) && it < S.ls_max;
    ++it;
    used = alpha;
    alpha *= (double)S.ls_tau;
  } while (cond);
  const bool accepted = it < S.ls_max;
  if (accepted) {
    const float a = (float)used;
    for (int i = v0 + tid; i < v1; i += nth)
#pragma unroll
      for (int k = 0; k < 3; ++k) S.x[3 * i + k] = S.x[3 * i + k] + a * (float)S.dx[3 * i + k];
    for (int e = e0 + tid; e < e1; e += nth)
#pragma unroll
      for (int k = 0; k < 6; ++k) S.s[6 * e + k] = S.s[6 * e + k] + a * S.ds[6 * e + k];
  }
  if (blockIdx.x == wg0 && threadIdx.x == 0) {
    if (newton_iter >= 0 && newton_iter < MFEM_MAX_STATS_ITERS) {
      double* st = S.stats + ((size_t)body * MFEM_MAX_STATS_ITERS + newton_iter) * MFEM_STATS_PER_ITER;
      st[2] = (double)it;
      st[3] = accepted ? 1.0 : 0.0;
      st[4] = used;
      st[5] = f0;
      st[6] = f1;
    }
    *ep_ptr = (int32_t)gs.epoch;
  }
}

__global__ void __launch_bounds__(kBlock) k_plain_update(mfem_solver S, int newton_iter) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  int i = blockIdx.x * kBlock + threadIdx.x;
  // (with self-contact, the CCD-bounded step: x + a dx, s + a ds)
  const float a = self_contact_on(S) ? S.sc_alpha[0] : 1.f;
  if (i < S.n_verts)
#pragma unroll
    for (int k = 0; k < 3; ++k) S.x[3 * i + k] += a * (float)S.dx[3 * i + k];
  if (i < S.n_tets)
#pragma unroll
    for (int k = 0; k < 6; ++k) S.s[6 * i + k] += a * S.ds[6 * i + k];
  if (i < (S.n_bodies > 1 ? S.n_bodies : 1) && newton_iter >= 0 && newton_iter < MFEM_MAX_STATS_ITERS) {
    double* st = S.stats + ((size_t)i * MFEM_MAX_STATS_ITERS + newton_iter) * MFEM_STATS_PER_ITER;
    st[2] = 0.0;
    st[3] = 1.0;
    st[4] = 1.0;
  }
}

__global__ void __launch_bounds__(kBlock) k_end_step(mfem_solver S, float dt) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // a grid barrier timed out: the step is void
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.n_verts) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) S.v[3 * i + k] = (S.x[3 * i + k] - S.x_start[3 * i + k]) / dt;
}

// Stand-alone SELL-64 SpMV y = H x: one workgroup per 64-row chunk, its four
// waves splitting the chunk's slots (wave w takes slots w, w + 4, ...: the
// loads of up to four slots per lane in flight at once, one dependent gather
// round trip per group), partial rows summed through LDS, and the chunk's 192
// outputs stored as three contiguous 256-byte runs.
__global__ void __launch_bounds__(kBlock) k_spmv(mfem_solver S, const float* xin, float* yout) {
  __shared__ float part[kWaves][3][64];
  const int c = blockIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int s0 = S.chunk_ptr[c], s1 = S.chunk_ptr[c + 1];
  float a0 = 0.f, a1 = 0.f, a2 = 0.f;
  constexpr int G = 4;
  for (int sb = s0 + w; sb < s1; sb += kWaves * G) {
    int col[G];
    float vv[G][9], g[G][3];
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const int sl = sb + kWaves * k;
      const bool ok = sl < s1;
      col[k] = ok ? S.sell_col[sl * 64 + lane] : 0;
      const float* vb = S.sell_val + (size_t)(ok ? sl : s0) * 576 + lane;
#pragma unroll
      for (int q = 0; q < 9; ++q) vv[k][q] = ok ? vb[64 * q] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < G; ++k)
#pragma unroll
      for (int a = 0; a < 3; ++a) g[k][a] = xin[3 * col[k] + a];
#pragma unroll
    for (int k = 0; k < G; ++k) {
      a0 += vv[k][0] * g[k][0] + vv[k][1] * g[k][1] + vv[k][2] * g[k][2];
      a1 += vv[k][3] * g[k][0] + vv[k][4] * g[k][1] + vv[k][5] * g[k][2];
      a2 += vv[k][6] * g[k][0] + vv[k][7] * g[k][1] + vv[k][8] * g[k][2];
    }
  }
  part[w][0][lane] = a0;
  part[w][1][lane] = a1;
  part[w][2][lane] = a2;
  __syncthreads();
  // thread t < 192 sums output t of the chunk (row t / 3, component t % 3)
  const int t = threadIdx.x;
  if (t < 192) {
    const int r = t / 3, a = t - 3 * r;
    const long row = (long)c * 64 + r;
    if (row < S.n_verts)
      yout[3 * row + a] = part[0][a][r] + part[1][a][r] + part[2][a][r] + part[3][a][r];
  }
}

__global__ void __launch_bounds__(kBlock) k_merit(mfem_solver S, float dt, float a, double* out) {
  __shared__ double sm[kWaves];
  const int tid = blockIdx.x * kBlock + threadIdx.x;
  double f = 0.0;
  if (tid < S.n_verts) f += vertex_merit(S, tid, a, dt);
  if (tid < S.n_tets) f += tet_merit(S, tid, a);
  if (tid < S.n_tris && tri_contact_on(S)) f += tri_merit(S, tid, a);
  if (tid < S.n_tris && shell_on(S)) f += shell_merit(S, tid, a);
  f = wave_sum(f);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = f;
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int w = 0; w < kWaves; ++w) acc += sm[w];
    atomicAdd(out, acc);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct PcgVariant {
  int cpw, rw, mode;  // mode 0 standard, 1 pipelined, 2 standard with four waves per SIMD, 3 pipelined with the coarse correction
  const void* fn;
};
#define PCG_V(C, R, P) {C, R, P, reinterpret_cast<const void*>(&k_pcg<C, R, P>)}
// resident-row variants first (one chunk per wave, rows in VGPRs), then the
// streaming ones (the low-register one holds twice the rows of the
// register-held one-chunk form in one grid) with more chunks per wave for
// larger systems.  Pipelined variants use an FP64 SpMV input (their FP64
// recurrences leave no room for resident rows); pipelining with an FP32 SpMV
// input stagnates near rel-res 1e-3 on the contact scenes and is not
// instantiated.
const PcgVariant kPcgVariants[] = {
    PCG_V(1, 16, 0), PCG_V(1, 0, 0), {1, 0, 2, reinterpret_cast<const void*>(&k_pcg_lowreg)},
    PCG_V(2, 0, 0), PCG_V(4, 0, 0), PCG_V(8, 0, 0),
    PCG_V(1, 16, 1), PCG_V(1, 0, 1), PCG_V(2, 0, 1), PCG_V(4, 0, 1), PCG_V(8, 0, 1),
    // rows of at most 15 blocks (valence <= 14 in a tetrahedral mesh, the
    // common case): no padding slot in every row's SpMV
    PCG_V(1, 15, 1),
    // block Jacobi plus the workgroup-local rigid-mode correction
    // (MFEM_PCG_COARSE): pipelined, resident rows
    PCG_V(1, 16, 3), PCG_V(1, 15, 3),
    // wider resident rows for refined meshes (a split edge raises the
    // valence of the vertices around it past 15)
    PCG_V(1, 20, 1), PCG_V(1, 24, 1), PCG_V(1, 20, 3),
};
#undef PCG_V
constexpr int kNumPcgVariants = sizeof(kPcgVariants) / sizeof(kPcgVariants[0]);

struct CoopInfo {
  int max_blocks_pcg[kNumPcgVariants] = {};
  int max_blocks_ls = 0;
  bool ready = false;
};
std::map<int, CoopInfo> g_coop;

int coop_info(CoopInfo** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  CoopInfo& ci = g_coop[dev];
  if (!ci.ready) {
    int cus = 0, coop = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) return fail(MFEM_EHIP, "device does not support cooperative launches");
    for (int k = 0; k < kNumPcgVariants; ++k) {
      int nb = 0;
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kPcgVariants[k].fn, kBlock, 0));
      ci.max_blocks_pcg[k] = nb * cus;
    }
    int nb = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_linesearch), kBlock, 0));
    ci.max_blocks_ls = nb * cus;
    ci.ready = true;
  }
  *out = &ci;
  return MFEM_OK;
}

int check_solver(const mfem_solver* S) {
  if (!S) return fail(MFEM_EINVAL, "null solver");
  if (S->n_verts <= 0 || S->n_tets <= 0) return fail(MFEM_EINVAL, "empty mesh");
  if (S->n_chunks != (S->n_verts + 63) / 64) return fail(MFEM_EINVAL, "n_chunks != ceil(n_verts/64)");
  if (S->n_slots < S->n_chunks) return fail(MFEM_EINVAL, "n_slots < n_chunks");
  if (S->max_width <= 0) return fail(MFEM_EINVAL, "max_width must be the widest chunk (slots)");
  if (S->n_shapes < 0 || S->n_tris < 0) return fail(MFEM_EINVAL, "negative count");
  if (S->iterations > MFEM_MAX_STATS_ITERS) return fail(MFEM_EINVAL, "iterations > MFEM_MAX_STATS_ITERS");
  if (S->local_sweeps < 0 || S->local_sweeps > 16 || (S->local_sweeps & 1) ||
      (S->local_sweeps > 0 && !(S->local_omega > 0.f)))
    return fail(MFEM_EINVAL, "local_sweeps must be even in [0, 16] and local_omega > 0");
  if (!S->x || !S->v || !S->s || !S->lam || !S->tets || !S->sell_val || !S->status || !S->stats)
    return fail(MFEM_EINVAL, "null buffer");
  if (S->n_bodies > 1) {
    if (!S->body_wg || !S->body_vert || !S->body_tet || !S->body_tri || !S->body_ls_wg || !S->body_epoch)
      return fail(MFEM_EINVAL, "a batch needs its body arrays");
    if (S->n_chunks % kWaves != 0) return fail(MFEM_EINVAL, "a batch needs whole workgroups of chunks per body");
    if (S->body_wg_align != 1 && S->body_wg_align != 2 && S->body_wg_align != 4 && S->body_wg_align != 8)
      return fail(MFEM_EINVAL, "body_wg_align must be 1, 2, 4 or 8");
    if ((S->n_chunks / kWaves) % S->body_wg_align != 0)
      return fail(MFEM_EINVAL, "a batch's blocks must be a multiple of body_wg_align");
    if (S->ls_blocks < S->n_bodies || S->ls_blocks > MFEM_RED_MAX_BLOCKS)
      return fail(MFEM_EINVAL, "ls_blocks must be in [n_bodies, MFEM_RED_MAX_BLOCKS]");
  }
  if (S->self_contact) {
    if (S->n_bodies > 1) return fail(MFEM_EINVAL, "self-contact needs a single body (n_bodies <= 1)");
    if (!S->sc_pt || !S->sc_ee || !S->sc_count || !S->sc_g || !S->sc_h || !S->sc_alpha || !S->sc_ws ||
        (S->n_sc_points > 0 && !S->sc_points) || (S->n_sc_edges > 0 && !S->sc_edges) || (S->n_tris > 0 && !S->tris))
      return fail(MFEM_EINVAL, "self-contact enabled without its buffers");
  }
  if (S->fused_local && !S->line_search)
    return fail(MFEM_EINVAL, "fused_local needs the line search (its first pass runs the local solve)");
  if (S->shell_on && S->n_tris > 0 &&
      (!S->tris || !S->tcon_ptr || !S->tcon_src || !S->vtri_ptr || !S->vtri_src || !S->rest || !S->tri_W ||
       !S->tri_area || !S->shell_h || !S->gS))
    return fail(MFEM_EINVAL, "membrane shell enabled without its buffers");
  if (S->tri_contact && S->n_tris > 0 &&
      (!S->tris || !S->tcon_ptr || !S->tcon_src || !S->vtri_ptr || !S->vtri_src || !S->tri_dist || !S->tri_bary ||
       !S->tri_nrm || !S->tri_sid))
    return fail(MFEM_EINVAL, "triangle contact enabled without its buffers");
  return MFEM_OK;
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MFEM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return MFEM_OK;
}

}  // namespace

#define FIELD_TABLE(X)                                                                                   \
  X(n_verts) X(n_tets) X(n_tris) X(n_shapes) X(n_chunks) X(n_slots) X(energy) X(iterations) X(cg_maxiter) \
  X(precond) X(line_search) X(ls_max) X(tri_contact) X(friction_on) X(cg_tol) X(ls_alpha0) X(ls_tau)      \
  X(ls_c) X(ls_thr) X(sing_thr) X(attach_k) X(d0) X(d1) X(kappa) X(eps_v) X(qb) X(gravity) X(max_width) X(pcg_flags) X(pcg_variant) X(tets) \
  X(dminv) X(materials) X(vol) X(mass) X(inv_mass) X(pin_pos) X(tris) X(shape_type) X(shape_xf)            \
  X(shape_scale) X(shape_mu) X(chunk_ptr) X(sell_col) X(sell_val) X(diag_lane) X(contrib_ptr)             \
  X(contrib_src) X(vinc_ptr) X(vinc_src) X(x) X(v) X(s) X(lam) X(x_tilde) X(x_start) X(lag_n) X(lag_fn)    \
  X(lag_sid) X(G) X(Hl) X(gl) X(Hsinv) X(gs) X(psi) X(Kb) X(fe) X(gx) X(gB) X(gD) X(Hv) X(dist) X(eB)      \
  X(eD) X(g) X(dx) X(ds) X(pz) X(px) X(red) X(flags) X(status) X(stats) X(mlp_w) X(mlp_wbf) X(ws_x) X(ws_aux) X(ws_y) X(guess) X(tcon_ptr) X(tcon_src) X(vtri_ptr) X(vtri_src) X(tri_dist) X(tri_bary) X(tri_nrm) X(tri_sid) X(n_bodies) X(ls_blocks) X(body_wg) X(body_vert) X(body_tet) X(body_tri) X(body_ls_wg) X(body_epoch) X(tet_energy) X(body_wg_align) X(local_sweeps) X(local_omega) X(dx_hist) X(dx_same) X(dx_same_iters) X(coarse_w) X(coarse_omega) X(cw_x) X(cw_y) X(cw_w) X(cw_wbf) X(self_contact) X(n_sc_points) X(n_sc_edges) X(sc_max_pairs) X(sc_max_span) X(sc_dhat) X(sc_kappa) X(sc_cell) X(sc_ccd_s) X(sc_points) X(sc_edges) X(sc_pt) X(sc_ee) X(sc_count) X(sc_g) X(sc_h) X(sc_alpha) X(sc_ws) X(sc_ws_bytes) X(shell_on) X(shell_mu) X(rest) X(tri_W) X(tri_area) X(shell_h) X(gS) X(fused_local)

extern "C" {

int64_t mfem_solver_sizeof(void) { return (int64_t)sizeof(mfem_solver); }

int64_t mfem_solver_offsetof(const char* field) {
#define X(f) \
  if (!strcmp(field, #f)) return (int64_t)offsetof(mfem_solver, f);
  FIELD_TABLE(X)
#undef X
  return -1;
}

int mfem_version(void) { return MFEM_VERSION; }

int mfem_wall_clock_khz(void) {
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return -1;
  return khz;
}

const char* mfem_last_error(void) { return g_last_error.c_str(); }

int mfem_init_stretch(const mfem_solver* S, void* stream) {
  if (int rc = check_solver(S)) return rc;
  hipLaunchKernelGGL(k_init_stretch, dim3(blocks_for(S->n_tets)), dim3(kBlock), 0, (hipStream_t)stream, *S);
  return launched();
}

int mfem_begin_step(const mfem_solver* S, float dt, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!(dt > 0.f)) return fail(MFEM_EINVAL, "dt must be > 0");
  hipLaunchKernelGGL(k_begin_step, dim3(blocks_for(S->n_verts)), dim3(kBlock), 0, (hipStream_t)stream, *S, dt);
  if (int rc = launched()) return rc;
  // self-contact: candidates on the static boxes of the step's start
  if (S->self_contact)
    if (int rc = mfem_self_contact_candidates(S, 0, stream))
      return fail(rc, std::string("self-contact: ") + mfem_self_contact_last_error());
  return MFEM_OK;
}

int mfem_update_system(const mfem_solver* S, float dt, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!(dt > 0.f)) return fail(MFEM_EINVAL, "dt must be > 0");
  hipStream_t st = (hipStream_t)stream;
  if (S->tri_contact && S->n_tris > 0 && S->n_shapes > 0) {
    hipLaunchKernelGGL(k_tri_contact, dim3(blocks_for(S->n_tris)), dim3(kBlock), 0, st, *S);
    if (int rc = launched()) return rc;
  }
  if (S->self_contact)
    if (int rc = mfem_self_contact_eval(S, stream))
      return fail(rc, std::string("self-contact: ") + mfem_self_contact_last_error());
  // one launch: the tets' element work, then the vertices' terms
  const int nt = (int)blocks_for(S->n_tets), nv = (int)blocks_for(S->n_verts);
  if (S->fused_local)
    hipLaunchKernelGGL(k_element<false>, dim3(nt + nv), dim3(kBlock), 0, st, *S, dt, nt);
  else
    hipLaunchKernelGGL(k_element<true>, dim3(nt + nv), dim3(kBlock), 0, st, *S, dt, nt);
  return launched();
}

int mfem_shell_refresh(const mfem_solver* S, void* stream) {
  // (needs the topology and the shell's buffers only: state need not be bound yet)
  if (!S) return fail(MFEM_EINVAL, "null solver");
  if (!S->shell_on || S->n_tris == 0) return MFEM_OK;
  if (S->n_tris < 0 || S->n_slots <= 0) return fail(MFEM_EINVAL, "bad sizes");
  if (!(S->shell_mu > 0.f)) return fail(MFEM_EINVAL, "shell_mu must be > 0 with the shell on");
  if (!S->tris || !S->tcon_ptr || !S->tcon_src || !S->rest || !S->tri_W || !S->tri_area || !S->shell_h)
    return fail(MFEM_EINVAL, "membrane shell enabled without its buffers");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_shell_frames, dim3(blocks_for(S->n_tris)), dim3(kBlock), 0, st, *S);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(k_shell_hess, dim3(blocks_for((long)S->n_slots * 64)), dim3(kBlock), 0, st, *S);
  return launched();
}

int mfem_assemble(const mfem_solver* S, float dt, void* stream) {
  if (int rc = check_solver(S)) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_assemble, dim3(blocks_for((long)S->n_slots * 64)), dim3(kBlock), 0, st, *S, dt);
  return launched();
}

int mfem_global_solve(const mfem_solver* S, int newton_iter, void* stream) {
  if (int rc = check_solver(S)) return rc;
  CoopInfo* ci = nullptr;
  if (int rc = coop_info(&ci)) return rc;
  int sel = -1;
  unsigned nblocks = 0;
  if (S->pcg_variant >= 0) {  // forced (autotuner)
    int32_t info[5];
    if (int rc = mfem_pcg_variant_info(S, S->pcg_variant, info)) return rc;
    if (!info[4]) return fail(MFEM_EINVAL, "pcg_variant does not apply to this system");
    sel = S->pcg_variant;
    nblocks = (unsigned)info[3];
  }
  for (int k0 = 0; k0 < kNumPcgVariants && sel < 0; ++k0) {
    // resident rows: the narrowest form of this kind that holds the widest
    // chunk stands in for the others of its kind
    int k = k0;
    if (kPcgVariants[k0].rw > 0) {
      int best = -1;
      for (int j = 0; j < kNumPcgVariants; ++j) {
        const PcgVariant& u = kPcgVariants[j];
        if (u.rw > 0 && u.mode == kPcgVariants[k0].mode && u.cpw == kPcgVariants[k0].cpw && u.rw >= S->max_width &&
            (best < 0 || u.rw < kPcgVariants[best].rw))
          best = j;
      }
      if (best >= 0) k = best;
    }
    const PcgVariant& v = kPcgVariants[k];
    const bool want = (S->pcg_flags & MFEM_PCG_PIPELINED) != 0;
    const bool coarse = (S->pcg_flags & MFEM_PCG_COARSE) != 0;
    if ((v.mode == 3) != coarse) continue;  // the coarse correction: its own variants, and only they
    if (!coarse && (v.mode == 1) != want) continue;
    if ((v.mode == 1 || v.mode == 3) && S->local_sweeps > 0) continue;  // the local preconditioner is standard-mode only
    if (S->n_bodies > 1 && S->body_wg_align % v.cpw != 0) continue;  // a batch: body-aligned workgroups
    if (v.rw > 0 && ((S->pcg_flags & MFEM_PCG_STREAM_ROWS) || S->max_width > v.rw)) continue;
    long waves = (S->n_chunks + v.cpw - 1) / v.cpw;
    long nb = (waves + kWaves - 1) / kWaves;
    if (nb <= ci->max_blocks_pcg[k] && nb <= MFEM_RED_MAX_BLOCKS) {
      sel = k;
      nblocks = (unsigned)nb;
      break;
    }
  }
  if (sel < 0) return fail(MFEM_ECAPACITY, "system too large for the persistent PCG kernel");
  g_last_pcg_variant = sel;
  mfem_solver Sc = *S;
  void* args[] = {(void*)&Sc, (void*)&newton_iter};
  if (S->pcg_flags & MFEM_PCG_COOPERATIVE)
    HIP_TRY(hipLaunchCooperativeKernel(kPcgVariants[sel].fn, dim3(nblocks), dim3(kBlock), args, 0, (hipStream_t)stream));
  else
    HIP_TRY(hipLaunchKernel(kPcgVariants[sel].fn, dim3(nblocks), dim3(kBlock), args, 0, (hipStream_t)stream));
  return launched();
}

int mfem_local_solve(const mfem_solver* S, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (S->fused_local) return MFEM_OK;  // runs inside mfem_line_search
  hipLaunchKernelGGL(k_local, dim3(blocks_for(S->n_tets)), dim3(kBlock), 0, (hipStream_t)stream, *S);
  return launched();
}

int mfem_line_search(const mfem_solver* S, float dt, int newton_iter, void* stream) {
  if (int rc = check_solver(S)) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (S->self_contact) {
    // candidates on the boxes swept by the Newton step (every pair within
    // dhat anywhere on x + a dx, a in [0, 1]: the merit's and the next
    // iteration's), then the CCD bound of the first step
    if (int rc = mfem_self_contact_candidates(S, 1, stream))
      return fail(rc, std::string("self-contact: ") + mfem_self_contact_last_error());
    if (int rc = mfem_self_contact_ccd(S, stream))
      return fail(rc, std::string("self-contact: ") + mfem_self_contact_last_error());
  }
  if (!S->line_search) {
    long n = S->n_verts > S->n_tets ? S->n_verts : S->n_tets;
    hipLaunchKernelGGL(k_plain_update, dim3(blocks_for(n)), dim3(kBlock), 0, st, *S, newton_iter);
    return launched();
  }
  CoopInfo* ci = nullptr;
  if (int rc = coop_info(&ci)) return rc;
  long n = S->n_verts > S->n_tets ? S->n_verts : S->n_tets;
  long nb = blocks_for(n);
  if (nb > ci->max_blocks_ls) nb = ci->max_blocks_ls;
  if (nb > MFEM_RED_MAX_BLOCKS) nb = MFEM_RED_MAX_BLOCKS;
  if (S->n_bodies > 1) {
    if (S->ls_blocks > ci->max_blocks_ls) return fail(MFEM_ECAPACITY, "ls_blocks exceeds the resident workgroups");
    nb = S->ls_blocks;
  }
  if (nb < 1) return fail(MFEM_EHIP, "line-search kernel cannot be resident");
  mfem_solver Sc = *S;
  void* args[] = {(void*)&Sc, (void*)&dt, (void*)&newton_iter};
  if (S->pcg_flags & MFEM_PCG_COOPERATIVE)
    HIP_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_linesearch), dim3((unsigned)nb), dim3(kBlock),
                                       args, 0, st));
  else
    HIP_TRY(hipLaunchKernel(reinterpret_cast<const void*>(&k_linesearch), dim3((unsigned)nb), dim3(kBlock), args, 0,
                            st));
  return launched();
}

int mfem_end_step(const mfem_solver* S, float dt, void* stream) {
  if (int rc = check_solver(S)) return rc;
  hipLaunchKernelGGL(k_end_step, dim3(blocks_for(S->n_verts)), dim3(kBlock), 0, (hipStream_t)stream, *S, dt);
  return launched();
}

int mfem_step(const mfem_solver* S, float dt, void* stream) {
  if (int rc = mfem_begin_step(S, dt, stream)) return rc;
  for (int it = 0; it < S->iterations; ++it) {
    if (int rc = mfem_update_system(S, dt, stream)) return rc;
    if (int rc = mfem_assemble(S, dt, stream)) return rc;
    if ((S->pcg_flags & MFEM_PCG_LEARNED_GUESS) && S->guess && S->mlp_wbf && S->mlp_w) {
      if (int rc = mfem_ws_features(S, dt, it, stream)) return rc;
      if (int rc = mfem_mlp_forward(S->ws_x, S->n_verts, S->mlp_wbf, S->mlp_w, S->ws_y, stream))
        return fail(rc, std::string("mfem_mlp_forward: ") + mfem_mlp_last_error());
      if (int rc = mfem_ws_apply(S, stream)) return rc;
    }
    if (it == 0 && (S->pcg_flags & MFEM_PCG_COARSE) && S->cw_wbf && S->cw_w && S->coarse_w) {
      // the coarse-weight network, once per step (its features come from
      // the first Newton iteration's system)
      if (int rc = mfem_coarse_features(S, dt, stream)) return rc;
      if (int rc = mfem_mlp_forward(S->cw_x, S->n_verts, S->cw_wbf, S->cw_w, S->cw_y, stream))
        return fail(rc, std::string("mfem_mlp_forward: ") + mfem_mlp_last_error());
      if (int rc = mfem_coarse_weights(S, stream)) return rc;
    }
    if (int rc = mfem_global_solve(S, it, stream)) return rc;
    if (int rc = mfem_local_solve(S, stream)) return rc;
    if (int rc = mfem_line_search(S, dt, it, stream)) return rc;
  }
  return mfem_end_step(S, dt, stream);
}

int mfem_pcg_variant_count(void) { return kNumPcgVariants; }

int mfem_last_pcg_variant(void) { return g_last_pcg_variant; }

int mfem_line_search_max_blocks(void) {
  CoopInfo* ci = nullptr;
  if (coop_info(&ci)) return 0;
  return ci->max_blocks_ls < MFEM_RED_MAX_BLOCKS ? ci->max_blocks_ls : MFEM_RED_MAX_BLOCKS;
}

int mfem_pcg_variant_info(const mfem_solver* S, int k, int32_t* info) {
  if (!S || !info || k < 0 || k >= kNumPcgVariants) return fail(MFEM_EINVAL, "bad PCG variant query");
  CoopInfo* ci = nullptr;
  if (int rc = coop_info(&ci)) return rc;
  const PcgVariant& v = kPcgVariants[k];
  const long waves = (S->n_chunks + v.cpw - 1) / v.cpw;
  const long nb = (waves + kWaves - 1) / kWaves;
  info[0] = v.cpw;
  info[1] = v.rw;
  info[2] = v.mode;
  info[3] = (int32_t)nb;
  info[4] = (nb <= ci->max_blocks_pcg[k] && nb <= MFEM_RED_MAX_BLOCKS && (v.rw == 0 || S->max_width <= v.rw) &&
             (S->n_bodies <= 1 || S->body_wg_align % v.cpw == 0) &&
             !((v.mode == 1 || v.mode == 3) && S->local_sweeps > 0) &&
             (v.mode == 3) == ((S->pcg_flags & MFEM_PCG_COARSE) != 0))
                ? 1
                : 0;
  return MFEM_OK;
}

int mfem_ws_features(const mfem_solver* S, float dt, int newton_iter, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!S->ws_x || !S->ws_aux) return fail(MFEM_EINVAL, "warm-start buffers not set");
  hipLaunchKernelGGL(k_ws_features, dim3(blocks_for(S->n_verts)), dim3(kBlock), 0, (hipStream_t)stream, *S, dt,
                     newton_iter);
  return launched();
}

int mfem_coarse_features(const mfem_solver* S, float dt, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!S->cw_x) return fail(MFEM_EINVAL, "mfem_coarse_features: cw_x is null");
  if (!(dt > 0.f)) return fail(MFEM_EINVAL, "dt must be > 0");
  hipLaunchKernelGGL(k_coarse_features, dim3(blocks_for(S->n_verts)), dim3(kBlock), 0, (hipStream_t)stream, *S, dt);
  return launched();
}

int mfem_coarse_weights(const mfem_solver* S, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!S->cw_y || !S->coarse_w) return fail(MFEM_EINVAL, "mfem_coarse_weights: cw_y / coarse_w is null");
  hipLaunchKernelGGL(k_coarse_weights, dim3(blocks_for(S->n_verts)), dim3(kBlock), 0, (hipStream_t)stream, *S);
  return launched();
}

int mfem_ws_apply(const mfem_solver* S, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!S->ws_aux || !S->ws_y || !S->guess) return fail(MFEM_EINVAL, "warm-start buffers not set");
  hipLaunchKernelGGL(k_ws_apply, dim3(blocks_for(S->n_verts)), dim3(kBlock), 0, (hipStream_t)stream, *S);
  return launched();
}

int mfem_spmv(const mfem_solver* S, const float* x, float* y, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!x || !y) return fail(MFEM_EINVAL, "null vector");
  hipLaunchKernelGGL(k_spmv, dim3(S->n_chunks), dim3(kBlock), 0, (hipStream_t)stream, *S, x, y);
  return launched();
}

__global__ void __launch_bounds__(kBlock) k_vertex_distance(mfem_solver S, float* vdist, int32_t* vsid) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= S.n_verts) return;
  const float x[3] = {S.x[3 * i], S.x[3 * i + 1], S.x[3 * i + 2]};
  float grad[3];
  int sid = -1;
  double d = 1.0e8;
  if (S.n_shapes > 0) shape_distance(S, x, grad, &sid, &d);
  vdist[i] = (float)d;
  vsid[i] = sid;
}

// world pose of each shape: body_q[shape_body[s]] * local[s] for a shape on a
// body, local[s] for a static one (shape_body < 0); poses are p(3), q(4) xyzw
__global__ void k_pose_shapes(const float* body_q, int n_body, const int32_t* shape_body, const float* local,
                              int n_shapes, float* shape_xf) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_shapes) return;
  const float* l = local + 7 * s;
  const int b = shape_body[s];
  float o[7];
  if (b < 0 || b >= n_body) {
#pragma unroll
    for (int k = 0; k < 7; ++k) o[k] = l[k];
  } else {
    const float* t = body_q + 7 * b;
    float R[9];
    quat_to_mat(t + 3, R);
#pragma unroll
    for (int i = 0; i < 3; ++i) o[i] = t[i] + R[3 * i] * l[0] + R[3 * i + 1] * l[1] + R[3 * i + 2] * l[2];
    // Hamilton product q_body * q_local (apply the local rotation first)
    const float x1 = t[3], y1 = t[4], z1 = t[5], w1 = t[6];
    const float x2 = l[3], y2 = l[4], z2 = l[5], w2 = l[6];
    o[3] = w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2;
    o[4] = w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2;
    o[5] = w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2;
    o[6] = w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2;
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) shape_xf[7 * s + k] = o[k];
}

int mfem_pose_shapes(const float* body_q, int n_body, const int32_t* shape_body, const float* shape_local,
                     int n_shapes, float* shape_xf, void* stream) {
  if (n_shapes < 0 || n_body < 0) return fail(MFEM_EINVAL, "mfem_pose_shapes: negative count");
  if (n_shapes == 0) return MFEM_OK;
  if (!shape_body || !shape_local || !shape_xf || (n_body > 0 && !body_q))
    return fail(MFEM_EINVAL, "mfem_pose_shapes: null pointer");
  hipLaunchKernelGGL(k_pose_shapes, dim3((n_shapes + 63) / 64), dim3(64), 0, (hipStream_t)stream, body_q, n_body,
                     shape_body, shape_local, n_shapes, shape_xf);
  HIP_TRY(hipGetLastError());
  return MFEM_OK;
}

int mfem_contact_distance(const mfem_solver* S, float* vdist, int32_t* vsid, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!vdist || !vsid) return fail(MFEM_EINVAL, "null output");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_vertex_distance, dim3(blocks_for(S->n_verts)), dim3(kBlock), 0, st, *S, vdist, vsid);
  if (int rc = launched()) return rc;
  if (S->tri_contact && S->n_tris > 0 && S->n_shapes > 0) {
    hipLaunchKernelGGL(k_tri_contact, dim3(blocks_for(S->n_tris)), dim3(kBlock), 0, st, *S);
    if (int rc = launched()) return rc;
  }
  return MFEM_OK;
}

int mfem_merit(const mfem_solver* S, float dt, float alpha, double* out, void* stream) {
  if (int rc = check_solver(S)) return rc;
  if (!out) return fail(MFEM_EINVAL, "null output");
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipMemsetAsync(out, 0, sizeof(double), st));
  long n = S->n_verts > S->n_tets ? S->n_verts : S->n_tets;
  if (S->n_tris > n) n = S->n_tris;
  hipLaunchKernelGGL(k_merit, dim3(blocks_for(n)), dim3(kBlock), 0, st, *S, dt, alpha, out);
  return launched();
}

}  // extern "C"
