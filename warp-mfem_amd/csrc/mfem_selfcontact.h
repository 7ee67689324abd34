// mfem_selfcontact.h -- device functions of the self-contact barrier shared
// by its kernels (mfem_selfcontact.hip) and the line search
// (mfem_kernels.hip).
//
// A candidate pair is a point-triangle pair (vertex v, surface triangle t)
// or an edge-edge pair (surface edges e1 < e2).  Its distance is that of
// the closest points (Ericson; src/mfem/collision.py:220-322), formed in FP64
// from the FP32 positions, and its energy is the IPC barrier
//   b(d) = -(d - dhat)^2 ln(d / dhat),  0 < d < dhat
// (src/mfem/collision.py:529-546) times kappa.  The gradient with respect to
// the pair's vertex j is kappa b'(d) cw_j n, n the unit vector between the
// closest points, cw_j the vertex's weight in the closest-point difference:
// (1, -w0, -w1, -w2) for the point and the triangle's barycentric weights,
// (1 - s, s, -(1 - t), -t) for the edges' parameters (the derivative of the
// distance with the closest-point parameters held, which is the gradient
// where the closest points are unique).
#pragma once

#include "../../include/mfem_amd.h"
#include "mfem_math.h"

namespace mfem {

// FP64 IPC barrier without extension: energy and its first two derivatives
MFEM_DEV void ipc_barrier64(double d, double dh, double& e, double& de, double& d2e) {
  const double l = log(d / dh), u = d - dh;
  e = -u * u * l;
  de = -(2.0 * u * l + u * u / d);
  d2e = -(2.0 * l + 4.0 * u / d - u * u / (d * d));
}

// vertex ids of candidate pair k (KIND 0: point-triangle, 1: edge-edge)
template <int KIND>
MFEM_DEV void sc_pair_vertices(const mfem_solver& S, int k, int* v) {
  const int32_t* P = KIND == 0 ? S.sc_pt : S.sc_ee;
  const int p0 = P[2 * k], p1 = P[2 * k + 1];
  if (KIND == 0) {
    v[0] = p0;
    v[1] = S.tris[3 * p1];
    v[2] = S.tris[3 * p1 + 1];
    v[3] = S.tris[3 * p1 + 2];
  } else {
    v[0] = S.sc_edges[2 * p0];
    v[1] = S.sc_edges[2 * p0 + 1];
    v[2] = S.sc_edges[2 * p1];
    v[3] = S.sc_edges[2 * p1 + 1];
  }
}

// distance of four points (FP64) as pair KIND, with the vertices' weights cw
// and the unit vector n (zero at zero distance)
template <int KIND>
MFEM_DEV double sc_distance(const double (&X)[4][3], double* cw, double* n) {
  double dv[3];
  if (KIND == 0) {
    double w[3];
    closest_point_triangle_bary<double>(X[0], X[1], X[2], X[3], w);
#pragma unroll
    for (int c = 0; c < 3; ++c) dv[c] = X[0][c] - (w[0] * X[1][c] + w[1] * X[2][c] + w[2] * X[3][c]);
    cw[0] = 1.0; cw[1] = -w[0]; cw[2] = -w[1]; cw[3] = -w[2];
  } else {
    double c1[3], c2[3], t;
    const double s = segment_segment<double>(X[0], X[1], X[2], X[3], c1, c2, &t);
#pragma unroll
    for (int c = 0; c < 3; ++c) dv[c] = c1[c] - c2[c];
    cw[0] = 1.0 - s; cw[1] = s; cw[2] = -(1.0 - t); cw[3] = -t;
  }
  const double d = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
  const double inv = d > 0.0 ? 1.0 / d : 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) n[c] = dv[c] * inv;
  return d;
}

// positions of pair k's vertices at x + a dx (the FP32 trial position of the
// line search, x itself at a = 0), widened to FP64
template <int KIND>
MFEM_DEV void sc_pair_positions(const mfem_solver& S, int k, float a, int* v, double (&X)[4][3]) {
  sc_pair_vertices<KIND>(S, k, v);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) X[j][c] = (double)(S.x[3 * v[j] + c] + a * (float)S.dx[3 * v[j] + c]);
}

// kappa b(d) of pair k at x + a dx; a pair at zero distance gives +1e300
// (the CCD keeps every accepted step away from it)
template <int KIND>
MFEM_DEV double sc_pair_merit(const mfem_solver& S, int k, float a) {
  int v[4];
  double X[4][3], cw[4], n[3];
  sc_pair_positions<KIND>(S, k, a, v, X);
  const double d = sc_distance<KIND>(X, cw, n);
  const double dh = S.sc_dhat;
  if (!(d < dh)) return 0.0;
  if (!(d > 0.0)) return 1.0e300;
  double e, de, d2e;
  ipc_barrier64(d, dh, e, de, d2e);
  return (double)S.sc_kappa * e;
}

MFEM_DEV int sc_pair_count(const mfem_solver& S, int kind) {
  const int n = S.sc_count[kind];
  return n < S.sc_max_pairs ? n : S.sc_max_pairs;
}

MFEM_DEV bool self_contact_on(const mfem_solver& S) { return S.self_contact != 0 && S.n_bodies <= 1; }

}  // namespace mfem
