// mfem_surface.hip -- point-to-surface queries of the surface-tracking and
// correspondence losses (src/mfem/refinement/surface_loss.py:202-245,
// 456-502), on the GPU in FP64.
//
// A query point's distance to a triangle mesh is taken the way the
// reference takes it: find its k nearest mesh vertices (cKDTree there, an
// exhaustive search here), then project it onto every triangle incident to
// those vertices (Ericson's closest point on a triangle, in the reference's
// region priority) and keep the smallest squared distance; when no
// incident triangle is usable, the nearest vertex's squared distance.
//
// k_knn: one wave per 64 query points; the candidate vertices stream
// through LDS 64 at a time (every lane reads the same LDS word: a
// broadcast) and each lane keeps its K best (distance, index) pairs in
// registers.  Candidates are visited in increasing index order and
// inserted on a strict improvement, so ties keep the lower index: the
// result is deterministic.  FP64 throughout, so near-ties resolve as an FP64
// KD-tree resolves them.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <string>

#include "../../include/mfem_amd.h"

namespace {

constexpr int kWave = 64;
thread_local std::string g_err;

int fail(int rc, const char* msg) {
  g_err = msg;
  return rc;
}
int launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MFEM_EHIP, hipGetErrorString(e));
  return MFEM_OK;
}
inline unsigned blocks(long n, int t) { return (unsigned)((n + t - 1) / t); }

template <int K>
__global__ void __launch_bounds__(kWave) k_knn(const double* __restrict__ q, int nq, const double* __restrict__ v,
                                              const int32_t* __restrict__ vmap, int m, int k,
                                              int32_t* __restrict__ out_idx, double* __restrict__ out_d2) {
  __shared__ double sx[kWave], sy[kWave], sz[kWave];
  __shared__ int32_t sid[kWave];
  const int i = blockIdx.x * kWave + threadIdx.x;
  const bool act = i < nq;
  const double qx = act ? q[3 * (size_t)i] : 0.0, qy = act ? q[3 * (size_t)i + 1] : 0.0,
               qz = act ? q[3 * (size_t)i + 2] : 0.0;
  double bd[K];
  int32_t bi[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    bd[r] = DBL_MAX;
    bi[r] = -1;
  }
  for (int base = 0; base < m; base += kWave) {
    const int j = base + threadIdx.x;
    if (j < m) {
      const int32_t g = vmap ? vmap[j] : j;
      sx[threadIdx.x] = v[3 * (size_t)g];
      sy[threadIdx.x] = v[3 * (size_t)g + 1];
      sz[threadIdx.x] = v[3 * (size_t)g + 2];
      sid[threadIdx.x] = g;
    }
    __syncthreads();
    const int cnt = min(kWave, m - base);
    for (int t = 0; t < cnt; ++t) {
      const double dx = qx - sx[t], dy = qy - sy[t], dz = qz - sz[t];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < bd[K - 1]) {
        // insertion into the sorted list (strict: an equal distance stays behind)
        double cd = d2;
        int32_t ci = sid[t];
#pragma unroll
        for (int r = 0; r < K; ++r) {
          if (cd < bd[r]) {
            const double td = bd[r];
            const int32_t ti = bi[r];
            bd[r] = cd;
            bi[r] = ci;
            cd = td;
            ci = ti;
          }
        }
      }
    }
    __syncthreads();
  }
  if (!act) return;
#pragma unroll
  for (int r = 0; r < K; ++r)
    if (r < k) {
      out_idx[(size_t)i * k + r] = bi[r];
      out_d2[(size_t)i * k + r] = bi[r] >= 0 ? bd[r] : DBL_MAX;
    }
}

__device__ __forceinline__ double dot3d(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ double safe_ratio(double num, double den) { return num / (den == 0.0 ? 1.0 : den); }

// closest point to p on triangle abc: Ericson's region tests, with the
// reference's priority (vertex A, B, C, then edge AB, AC, BC, then the
// face; surface_loss.py:47-93)
__device__ void closest_on_tri(const double* p, const double* a, const double* b, const double* c, double* out) {
  double ab[3], ac[3], ap[3], bp[3], cp[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ab[k] = b[k] - a[k];
    ac[k] = c[k] - a[k];
    ap[k] = p[k] - a[k];
    bp[k] = p[k] - b[k];
    cp[k] = p[k] - c[k];
  }
  const double d1 = dot3d(ab, ap), d2 = dot3d(ac, ap), d3 = dot3d(ab, bp), d4 = dot3d(ac, bp);
  const double d5 = dot3d(ab, cp), d6 = dot3d(ac, cp);
  const double va = d3 * d6 - d5 * d4, vb = d5 * d2 - d1 * d6, vc = d1 * d4 - d3 * d2;
  if (d1 <= 0.0 && d2 <= 0.0) {  // vertex A
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = a[k];
  } else if (d3 >= 0.0 && d4 <= d3) {  // vertex B
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = b[k];
  } else if (d6 >= 0.0 && d5 <= d6) {  // vertex C
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = c[k];
  } else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {  // edge AB
    const double t = safe_ratio(d1, d1 - d3);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = a[k] + ab[k] * t;
  } else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {  // edge AC
    const double t = safe_ratio(d2, d2 - d6);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = a[k] + ac[k] * t;
  } else if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0) {  // edge BC
    const double t = safe_ratio(d4 - d3, (d4 - d3) + (d5 - d6));
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = b[k] + (c[k] - b[k]) * t;
  } else {  // face
    const double den = va + vb + vc;
    const double inv = den == 0.0 ? 0.0 : 1.0 / den;
    const double vv = vb * inv, ww = vc * inv;
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = a[k] + ab[k] * vv + ac[k] * ww;
  }
}

// per query: smallest squared distance over the triangles incident to its
// k nearest vertices (in neighbour order, then incidence order; the first
// smallest wins), or the nearest vertex's when none is usable
__global__ void __launch_bounds__(256) k_project(const double* __restrict__ q, int nq, const double* __restrict__ v,
                                                 const int32_t* __restrict__ faces,
                                                 const int32_t* __restrict__ vert_faces, int deg,
                                                 const uint8_t* __restrict__ face_ok,
                                                 const int32_t* __restrict__ knn_idx,
                                                 const double* __restrict__ knn_d2, int k, double* __restrict__ out_sq,
                                                 int32_t* __restrict__ out_face, double* __restrict__ out_cp) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nq) return;
  const double p[3] = {q[3 * (size_t)i], q[3 * (size_t)i + 1], q[3 * (size_t)i + 2]};
  double best = DBL_MAX, bcp[3] = {0.0, 0.0, 0.0};
  int32_t bf = -1;
  for (int r = 0; r < k; ++r) {
    const int32_t n = knn_idx[(size_t)i * k + r];
    if (n < 0) continue;
    for (int s = 0; s < deg; ++s) {
      const int32_t f = vert_faces[(size_t)n * deg + s];
      if (f < 0) continue;
      if (face_ok && !face_ok[f]) continue;
      double a[3], b[3], c[3], cp[3];
      const int32_t* fv = faces + 3 * (size_t)f;
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        a[kk] = v[3 * (size_t)fv[0] + kk];
        b[kk] = v[3 * (size_t)fv[1] + kk];
        c[kk] = v[3 * (size_t)fv[2] + kk];
      }
      closest_on_tri(p, a, b, c, cp);
      const double dx = cp[0] - p[0], dy = cp[1] - p[1], dz = cp[2] - p[2];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < best) {
        best = d2;
        bf = f;
        bcp[0] = cp[0];
        bcp[1] = cp[1];
        bcp[2] = cp[2];
      }
    }
  }
  if (bf < 0) {
    const int32_t n = knn_idx[(size_t)i * k];
    best = n >= 0 ? knn_d2[(size_t)i * k] : DBL_MAX;
    if (n >= 0) {
      bcp[0] = v[3 * (size_t)n];
      bcp[1] = v[3 * (size_t)n + 1];
      bcp[2] = v[3 * (size_t)n + 2];
    }
  }
  out_sq[i] = best;
  if (out_face) out_face[i] = bf;
  if (out_cp) {
    out_cp[3 * (size_t)i] = bcp[0];
    out_cp[3 * (size_t)i + 1] = bcp[1];
    out_cp[3 * (size_t)i + 2] = bcp[2];
  }
}

}  // namespace

extern "C" {

const char* mfem_surface_last_error(void) { return g_err.c_str(); }

int mfem_surface_knn(const double* query, int n_query, const double* verts, const int32_t* vmap, int n_cand, int k,
                     int32_t* out_idx, double* out_d2, void* stream) {
  if (n_query < 0 || n_cand < 0) return fail(MFEM_EINVAL, "negative count");
  if (k < 1 || k > 8) return fail(MFEM_EINVAL, "k must be in [1, 8]");
  if (n_query == 0) return MFEM_OK;
  if (!query || !verts || !out_idx || !out_d2) return fail(MFEM_EINVAL, "null buffer");
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(blocks(n_query, kWave)), block(kWave);
  if (k == 1)
    hipLaunchKernelGGL(k_knn<1>, grid, block, 0, st, query, n_query, verts, vmap, n_cand, k, out_idx, out_d2);
  else if (k <= 2)
    hipLaunchKernelGGL(k_knn<2>, grid, block, 0, st, query, n_query, verts, vmap, n_cand, k, out_idx, out_d2);
  else if (k <= 4)
    hipLaunchKernelGGL(k_knn<4>, grid, block, 0, st, query, n_query, verts, vmap, n_cand, k, out_idx, out_d2);
  else
    hipLaunchKernelGGL(k_knn<8>, grid, block, 0, st, query, n_query, verts, vmap, n_cand, k, out_idx, out_d2);
  return launched();
}

int mfem_surface_project(const double* query, int n_query, const double* verts, const int32_t* faces,
                         const int32_t* vert_faces, int deg, const uint8_t* face_ok, const int32_t* knn_idx,
                         const double* knn_d2, int k, double* out_sq, int32_t* out_face, double* out_cp,
                         void* stream) {
  if (n_query < 0 || deg < 0 || k < 1) return fail(MFEM_EINVAL, "bad sizes");
  if (n_query == 0) return MFEM_OK;
  if (!query || !verts || !knn_idx || !knn_d2 || !out_sq || (deg > 0 && (!faces || !vert_faces)))
    return fail(MFEM_EINVAL, "null buffer");
  hipLaunchKernelGGL(k_project, dim3(blocks(n_query, 256)), dim3(256), 0, (hipStream_t)stream, query, n_query, verts,
                     faces, vert_faces, deg, face_ok, knn_idx, knn_d2, k, out_sq, out_face, out_cp);
  return launched();
}

}  // extern "C"
