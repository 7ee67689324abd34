// mfem_selfcontact.hip -- IPC point-triangle / edge-edge self-contact inside
// the Newton solve (gfx950): candidates, barrier gradient and Hessian, and
// the CCD bound on the line search's first step.
//
// Per step (mfem_step): candidates from the static boxes at the start;
// per Newton iteration: the pairs' gradient and diagonal Hessian blocks
// into sc_g / sc_h (added to g and H by k_assemble), the PCG, the local solve,
// candidates again on the boxes swept by the Newton step dx (so that they
// hold every pair within dhat anywhere on x + a dx, a in [0, 1], and at the
// next iteration's x), the CCD bound, then the line search with the pairs'
// energy in its merit.  The barrier terms are summed with FP64 atomics: the
// order of those sums varies run to run (last-bit differences in the FP64
// sums; the FP32 values added to g and H are almost always the same bits).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/mfem_amd.h"
#include "mfem_math.h"
#include "mfem_selfcontact.h"

using namespace mfem;

namespace {

constexpr int kThreads = 256;
// additive CCD: advancement steps per pair before the bound stops where it is
constexpr int kCcdMaxSteps = 256;

thread_local std::string g_err;

int fail(int rc, const std::string& msg) {
  g_err = msg;
  return rc;
}
int launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MFEM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return MFEM_OK;
}
inline unsigned blocks(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

__global__ void __launch_bounds__(kThreads) k_sc_clear(mfem_solver S) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= S.n_sc_points) return;
  const int v = S.sc_points[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) S.sc_g[3 * v + c] = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) S.sc_h[6 * v + q] = 0.0;
}

// gradient and Gauss-Newton diagonal blocks kappa max(b'', 0) cw_j^2 n n^T of
// every candidate pair below dhat at x; one thread per pair slot, as many
// slots as the list can hold (the count is on the device)
template <int KIND>
__global__ void __launch_bounds__(kThreads) k_sc_pairs(mfem_solver S) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;  // the step is void
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= sc_pair_count(S, KIND)) return;
  int v[4];
  double X[4][3], cw[4], n[3];
  sc_pair_positions<KIND>(S, k, 0.f, v, X);
  const double d = sc_distance<KIND>(X, cw, n);
  if (!(d < (double)S.sc_dhat) || !(d > 0.0)) return;
  double e, de, d2e;
  ipc_barrier64(d, S.sc_dhat, e, de, d2e);
  const double kap = S.sc_kappa, g = kap * de, h = kap * fmax(d2e, 0.0);
  const double nn[6] = {n[0] * n[0], n[1] * n[1], n[2] * n[2], n[0] * n[1], n[0] * n[2], n[1] * n[2]};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (cw[j] == 0.0) continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) atomicAdd(&S.sc_g[3 * v[j] + c], g * cw[j] * n[c]);
    const double hw = h * cw[j] * cw[j];
#pragma unroll
    for (int q = 0; q < 6; ++q) atomicAdd(&S.sc_h[6 * v[j] + q], hw * nn[q]);
  }
}

// candidate overflow or a primitive spanning too many cells: contacts
// would be missed, so the step is void (error bit MFEM_ERR_CONTACT)
__global__ void k_sc_check(mfem_solver S) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && (S.sc_count[2] | S.sc_count[3]) != 0)
    atomicOr(&S.status[MFEM_STATUS_ERRBITS], MFEM_ERR_CONTACT);
}

__global__ void k_sc_alpha_init(mfem_solver S) {
  if (threadIdx.x == 0 && blockIdx.x == 0) S.sc_alpha[0] = 1.0f;
}

// additive CCD (Li, Kaufman, Jiang 2021, "Codimensional Incremental
// Potential Contact", Alg. 1, thickness 0): advance the pair along dx by
// steps of 0.9 of its distance over the bound l_p of the relative motion (the
// distance cannot shrink by more than l_p per unit step), starting with
// (1 - s) d0 / l_p, until the distance drops below s d0 or the step passes 1.
// The returned step keeps the pair at a positive distance for every step
// size up to it.  All four displacements minus their mean (same distances,
// smaller bound).
template <int KIND>
__device__ double accd_toi(double (&X)[4][3], double (&P)[4][3], double s) {
  double m[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c] += 0.25 * P[j][c];
  double len[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
#pragma unroll
    for (int c = 0; c < 3; ++c) P[j][c] -= m[c];
    len[j] = sqrt(P[j][0] * P[j][0] + P[j][1] * P[j][1] + P[j][2] * P[j][2]);
  }
  const double lp = KIND == 0 ? len[0] + fmax(len[1], fmax(len[2], len[3])) : fmax(len[0], len[1]) + fmax(len[2], len[3]);
  if (!(lp > 0.0)) return 1.0;
  double cw[4], n[3];
  double d = sc_distance<KIND>(X, cw, n);
  if (!(d > 0.0)) return 0.0;
  const double g = s * d;
  double t = 0.0, tl = (1.0 - s) * d / lp;
  for (int it = 0; it < kCcdMaxSteps; ++it) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) X[j][c] += tl * P[j][c];
    d = sc_distance<KIND>(X, cw, n);
    if (t > 0.0 && d < g) return t;
    t += tl;
    if (t > 1.0) return 1.0;
    tl = 0.9 * d / lp;
  }
  return t;
}

template <int KIND>
__global__ void __launch_bounds__(kThreads) k_sc_ccd(mfem_solver S) {
  if (S.status[MFEM_STATUS_ERRBITS] != 0) return;
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= sc_pair_count(S, KIND)) return;
  int v[4];
  double X[4][3], P[4][3];
  sc_pair_positions<KIND>(S, k, 0.f, v, X);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) P[j][c] = (double)(float)S.dx[3 * v[j] + c];
  const double t = accd_toi<KIND>(X, P, (double)S.sc_ccd_s);
  if (t < 1.0) {
    // positive floats order as their bit patterns: an integer minimum is
    // the float minimum, and the same whatever order the pairs arrive in
    const float tf = (float)t;
    atomicMin(reinterpret_cast<int*>(S.sc_alpha), __float_as_int(tf > 0.f ? tf : 0.f));
  }
}

int check_sc(const mfem_solver* S) {
  if (!S) return fail(MFEM_EINVAL, "null solver");
  if (!S->self_contact) return fail(MFEM_EINVAL, "self-contact is off");
  if (S->n_bodies > 1) return fail(MFEM_EINVAL, "self-contact needs a single body (n_bodies <= 1)");
  if (S->n_sc_points < 0 || S->n_sc_edges < 0 || S->sc_max_pairs <= 0 || S->n_tris < 0)
    return fail(MFEM_EINVAL, "bad self-contact sizes");
  if (!(S->sc_dhat > 0.f) || !(S->sc_kappa >= 0.f) || !(S->sc_cell > 0.f) || !(S->sc_ccd_s > 0.f) ||
      !(S->sc_ccd_s < 1.f) || S->sc_max_span < 1 || S->sc_max_span > 8)
    return fail(MFEM_EINVAL, "bad self-contact parameters (dhat, kappa, cell, ccd_s, max_span)");
  if ((S->n_sc_points > 0 && !S->sc_points) || (S->n_sc_edges > 0 && !S->sc_edges) || (S->n_tris > 0 && !S->tris) ||
      !S->sc_pt || !S->sc_ee || !S->sc_count || !S->sc_g || !S->sc_h || !S->sc_alpha || !S->sc_ws || !S->x ||
      !S->dx || !S->status)
    return fail(MFEM_EINVAL, "self-contact buffer missing");
  return MFEM_OK;
}

}  // namespace

extern "C" {

const char* mfem_self_contact_last_error(void) { return g_err.c_str(); }

size_t mfem_self_contact_workspace_bytes(int n_points, int n_tris, int n_edges, int max_span) {
  const size_t a = mfem_broadphase_workspace_bytes(n_tris, n_points, max_span);
  const size_t b = mfem_broadphase_workspace_bytes(n_edges, n_edges, max_span);
  return a > b ? a : b;
}

int mfem_self_contact_candidates(const mfem_solver* S, int swept, void* stream) {
  if (int rc = check_sc(S)) return rc;
  const size_t ws = (size_t)S->sc_ws_bytes;
  const double* dx = swept ? S->dx : nullptr;
  // the two lists share the workspace: stream order runs them one after the other
  if (int rc = mfem_broadphase_swept(0, S->x, dx, S->sc_points, S->n_sc_points, S->tris, S->n_tris, S->sc_dhat,
                                     S->sc_cell, S->sc_max_span, S->sc_pt, S->sc_max_pairs, S->sc_count + 0,
                                     S->sc_count + 2, S->sc_ws, ws, stream))
    return fail(rc, std::string("point-triangle broad phase: ") + mfem_collide_last_error());
  if (int rc = mfem_broadphase_swept(1, S->x, dx, S->sc_edges, S->n_sc_edges, nullptr, 0, S->sc_dhat, S->sc_cell,
                                     S->sc_max_span, S->sc_ee, S->sc_max_pairs, S->sc_count + 1, S->sc_count + 3,
                                     S->sc_ws, ws, stream))
    return fail(rc, std::string("edge-edge broad phase: ") + mfem_collide_last_error());
  hipLaunchKernelGGL(k_sc_check, dim3(1), dim3(64), 0, (hipStream_t)stream, *S);
  return launched();
}

int mfem_self_contact_eval(const mfem_solver* S, void* stream) {
  if (int rc = check_sc(S)) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (S->n_sc_points > 0) hipLaunchKernelGGL(k_sc_clear, dim3(blocks(S->n_sc_points)), dim3(kThreads), 0, st, *S);
  hipLaunchKernelGGL(k_sc_pairs<0>, dim3(blocks(S->sc_max_pairs)), dim3(kThreads), 0, st, *S);
  hipLaunchKernelGGL(k_sc_pairs<1>, dim3(blocks(S->sc_max_pairs)), dim3(kThreads), 0, st, *S);
  return launched();
}

int mfem_self_contact_ccd(const mfem_solver* S, void* stream) {
  if (int rc = check_sc(S)) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_sc_alpha_init, dim3(1), dim3(64), 0, st, *S);
  hipLaunchKernelGGL(k_sc_ccd<0>, dim3(blocks(S->sc_max_pairs)), dim3(kThreads), 0, st, *S);
  hipLaunchKernelGGL(k_sc_ccd<1>, dim3(blocks(S->sc_max_pairs)), dim3(kThreads), 0, st, *S);
  return launched();
}

}  // extern "C"
