// mfem_refine.hip -- adaptive tet refinement, one pass (gfx950).
//
// Replaces refine() of src/mfem/refinement/refinement.py:914-1322 ("legacy"
// scoring) with a deterministic pass: the reference's selection is
// race-dependent, so the result is defined as the sequential greedy
// selection below and computed in parallel as iterated local maxima:
//
//  * edge table: open addressing on 64-bit edge keys (lower id << 32 | higher
//    id); per edge an integer score (int64 fixed point, order-independent
//    sums), a penetration flag, its incident-tet count, votes, state and the
//    lowest incident tet (owner of the new vertex);
//  * rounds: every unsplit tet votes for its best available candidate edge
//    (score, then key); an edge with the votes of all its tets is selected;
//    tets on selected edges are split and their other edges become
//    unavailable.  Rounds end when nothing is selected.  This is greedy
//    maximal selection by priority (an edge is taken iff it is a local
//    maximum among the available edges that share a tet with it), so the
//    result equals the sequential greedy selection;
//  * counts, prefix scans, scatter of tets / triangles / new vertices.
//
// Arithmetic that feeds the selection (scores, midpoints, interpolation) uses
// explicit round-to-nearest operations, so a sequential FP32 restatement
// reproduces it bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/mfem_amd.h"
#include "mfem_math.h"

using namespace mfem;

namespace {

constexpr int kThreads = 256;
constexpr double kScoreOne = 1099511627776.0;  // 2^40 (score == threshold)
constexpr long long kScoreCap = 1LL << 56;

thread_local std::string g_err;
int fail(int rc, const char* m) {
  g_err = m;
  return rc;
}
int launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MFEM_EHIP, hipGetErrorString(e));
  return MFEM_OK;
}
inline unsigned blocks(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

enum : int32_t { kAvailable = 0, kUnavailable = 1, kSelected = 2 };

__device__ __forceinline__ uint64_t ekey(int a, int b) {
  return a < b ? ((uint64_t)(uint32_t)a << 32) | (uint32_t)b : ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
}
__device__ __forceinline__ uint32_t ehash(uint64_t k, uint32_t mask) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  return (uint32_t)k & mask;
}
// slot of key (inserted when absent); -1 when the table is full
__device__ int table_insert(unsigned long long* keys, uint32_t mask, uint64_t k) {
  uint32_t h = ehash(k, mask);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const unsigned long long prev = atomicCAS(&keys[h], 0ull, (unsigned long long)k);
    if (prev == 0ull || prev == k) return (int)h;
    h = (h + 1) & mask;
  }
  return -1;
}
__device__ int table_find(const unsigned long long* keys, uint32_t mask, uint64_t k) {
  uint32_t h = ehash(k, mask);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const unsigned long long cur = keys[h];
    if (cur == k) return (int)h;
    if (cur == 0ull) return -1;
    h = (h + 1) & mask;
  }
  return -1;
}

struct Table {
  unsigned long long* keys;
  unsigned long long* score;  // int64 fixed point, score 2^40 == threshold
  int32_t* pen;
  int32_t* cnt;
  int32_t* votes;
  int32_t* state;
  int32_t* owner;
  int32_t* newvid;
  uint32_t mask;
};

struct Mesh {
  const int32_t* tets;
  int n_tets;
  const float* x;
  const float* inv_mass;
  const float* tet_score;
  const float* vertex_score;
  // shapes for the midpoint penetration test (world poses)
  const int32_t* shape_type;
  const float* shape_xf;
  const float* shape_scale;
  int n_shapes;
};

// signed distance of p to the nearest shape (plane: local z; capsule: axis
// distance minus radius), src/mfem/refinement/contact.py:62-95
__device__ float min_signed_distance(const Mesh& M, const float* p) {
  float dm = 1.0e8f;
  for (int k = 0; k < M.n_shapes; ++k) {
    const float* xf = M.shape_xf + 7 * k;
    float R[9];
    quat_to_mat(xf + 3, R);
    const float dw[3] = {p[0] - xf[0], p[1] - xf[1], p[2] - xf[2]};
    float xl[3];
    for (int i = 0; i < 3; ++i) xl[i] = R[i] * dw[0] + R[3 + i] * dw[1] + R[6 + i] * dw[2];
    float d;
    if (M.shape_type[k] == MFEM_SHAPE_PLANE) {
      d = xl[2];
    } else if (M.shape_type[k] == MFEM_SHAPE_CAPSULE) {
      const float rad = M.shape_scale[3 * k], hh = M.shape_scale[3 * k + 1];
      const float az = fminf(fmaxf(xl[2], -hh), hh);
      d = sqrtf(xl[0] * xl[0] + xl[1] * xl[1] + (xl[2] - az) * (xl[2] - az)) - rad;
    } else {
      continue;
    }
    dm = fminf(dm, d);
  }
  return dm;
}

__device__ __forceinline__ long long to_fixed(float c, double q) {
  long long v = __double2ll_rn((double)c * q);
  return v < 0 ? 0 : (v > kScoreCap ? kScoreCap : v);
}

__global__ void k_fill(int32_t* p, long n, int32_t v) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill64(unsigned long long* p, long n) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// scores (populate_candidates, refinement.py:74-154)
__global__ void __launch_bounds__(kThreads) k_score(Mesh M, Table T, float w_tet, float w_vert, double q,
                                                    int32_t* status) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= M.n_tets) return;
  const int32_t* tv = M.tets + 4 * t;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j) {
      const int a = tv[i], b = tv[j];
      if (M.inv_mass[a] == 0.f && M.inv_mass[b] == 0.f) continue;
      const int s = table_insert(T.keys, T.mask, ekey(a, b));
      if (s < 0) {
        atomicOr(status, MFEM_REFINE_TABLE_FULL);
        return;
      }
      atomicAdd(&T.cnt[s], 1);
      atomicMin(&T.owner[s], t);
      const float xa[3] = {M.x[3 * a], M.x[3 * a + 1], M.x[3 * a + 2]};
      const float xb[3] = {M.x[3 * b], M.x[3 * b + 1], M.x[3 * b + 2]};
      if (M.n_shapes > 0) {
        const float mid[3] = {__fmul_rn(0.5f, __fadd_rn(xa[0], xb[0])), __fmul_rn(0.5f, __fadd_rn(xa[1], xb[1])),
                              __fmul_rn(0.5f, __fadd_rn(xa[2], xb[2]))};
        if (min_signed_distance(M, mid) < 0.f) {
          atomicOr(&T.pen[s], 1);
          continue;
        }
      }
      const float d0 = __fsub_rn(xb[0], xa[0]), d1 = __fsub_rn(xb[1], xa[1]), d2 = __fsub_rn(xb[2], xa[2]);
      const float L = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(d0, d0), __fmul_rn(d1, d1)), __fmul_rn(d2, d2)));
      const float vs = fminf(M.vertex_score[a], M.vertex_score[b]);
      const float c = __fmul_rn(__fadd_rn(__fmul_rn(M.tet_score[t], w_tet), __fmul_rn(vs, w_vert)), L);
      atomicAdd(&T.score[s], (unsigned long long)to_fixed(c, q));
    }
}

// --- "geometric" scoring (populate_candidates_geometric /
// populate_tri_candidates_geometric / elastic_density_stats,
// refinement.py:123-317).  Scores are maxima, so the order of the atomics
// does not matter; the arithmetic is FP64 with explicit round-to-nearest
// operations (no contraction), converted once to the fixed-point priority.

// strain-energy density over mu of tet t (0 for degenerate data)
__device__ double tet_density(const float* tet_score, const float* materials, const float* dminv, int t) {
  const float* m = dminv + 9 * t;
  // det of DmInv (row-major), cofactor expansion along the first row
  const double c0 = __dsub_rn(__dmul_rn(m[4], m[8]), __dmul_rn(m[5], m[7]));
  const double c1 = __dsub_rn(__dmul_rn(m[3], m[8]), __dmul_rn(m[5], m[6]));
  const double c2 = __dsub_rn(__dmul_rn(m[3], m[7]), __dmul_rn(m[4], m[6]));
  const double det = __dadd_rn(__dsub_rn(__dmul_rn(m[0], c0), __dmul_rn(m[1], c1)), __dmul_rn(m[2], c2));
  const double vol = fabs(__ddiv_rn(1.0, __dmul_rn(6.0, det)));
  const double mu = materials[3 * t];
  if (mu > 0.0 && vol > 0.0 && isfinite(vol)) {
    const double e = tet_score[t] > 0.f ? (double)tet_score[t] : 0.0;
    return __ddiv_rn(e, __dmul_rn(vol, mu));
  }
  return 0.0;
}

__device__ __forceinline__ double edge_len(const float* x, int a, int b) {
  const double d0 = __dsub_rn((double)x[3 * b], (double)x[3 * a]);
  const double d1 = __dsub_rn((double)x[3 * b + 1], (double)x[3 * a + 1]);
  const double d2 = __dsub_rn((double)x[3 * b + 2], (double)x[3 * a + 2]);
  return __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)), __dmul_rn(d2, d2)));
}

__device__ __forceinline__ double proximity(float d, double d1) {
  const double p = __ddiv_rn(__dsub_rn(d1, (double)d), d1);
  return p > 0.0 ? p : 0.0;
}

// per-block FP64 sums of the densities, in a fixed tree order
__global__ void __launch_bounds__(kThreads) k_density_partial(const float* tet_score, const float* materials,
                                                              const float* dminv, int n_tets, double* part) {
  __shared__ double sh[kThreads];
  const int t = blockIdx.x * kThreads + threadIdx.x;
  sh[threadIdx.x] = t < n_tets ? tet_density(tet_score, materials, dminv, t) : 0.0;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = __dadd_rn(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// total = ordered sum of the block sums (one workgroup)
__global__ void __launch_bounds__(kThreads) k_density_total(const double* part, int n_part, double* total) {
  __shared__ double sh[kThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n_part; i += kThreads) acc = __dadd_rn(acc, part[i]);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = __dadd_rn(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = sh[0];
}

struct Geo {
  const float* materials;
  const float* dminv;
  const float* vdist;
  const int32_t* vsid;
  const double* dens_total;  // sum of the densities over the active tets
  double h_min, d1, w_e, w_c, w_t, q;
};

__device__ __forceinline__ void score_max(Table T, int s, double c, double q) {
  double v = __dmul_rn(c, q);
  long long f = v > 0.0 ? __double2ll_rn(v) : 0;
  if (f > kScoreCap) f = kScoreCap;
  atomicMax(&T.score[s], (unsigned long long)f);
}

__global__ void __launch_bounds__(kThreads) k_score_geometric(Mesh M, Table T, Geo G, int32_t* status) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= M.n_tets) return;
  const int32_t* tv = M.tets + 4 * t;
  double s_el = tet_density(M.tet_score, G.materials, G.dminv, t);
  const double total = *G.dens_total;
  if (total > 0.0 && M.n_tets > 0) {
    const double mean = __ddiv_rn(total, (double)M.n_tets);
    const double ex = __dsub_rn(__ddiv_rn(s_el, mean), 1.0);
    s_el = ex > 0.0 ? ex : 0.0;
  }
  double l_max = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j) l_max = fmax(l_max, edge_len(M.x, tv[i], tv[j]));
  l_max = fmax(l_max, 1.0e-20);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j) {
      const int a = tv[i], b = tv[j];
      if (M.inv_mass[a] == 0.f && M.inv_mass[b] == 0.f) continue;
      const int s = table_insert(T.keys, T.mask, ekey(a, b));
      if (s < 0) {
        atomicOr(status, MFEM_REFINE_TABLE_FULL);
        return;
      }
      atomicAdd(&T.cnt[s], 1);
      atomicMin(&T.owner[s], t);
      const double L = edge_len(M.x, a, b);
      if (L < __dmul_rn(2.0, G.h_min)) continue;
      double s_c = 0.0;
      const int ends[2] = {a, b};
      for (int k = 0; k < 2; ++k) {
        const int sid = G.vsid ? G.vsid[ends[k]] : -1;
        if (sid >= 0 && sid < M.n_shapes && M.shape_type[sid] == MFEM_SHAPE_CAPSULE)
          s_c = fmax(s_c, proximity(G.vdist[ends[k]], G.d1));
      }
      const double c = __dmul_rn(__dmul_rn(__ddiv_rn(L, G.h_min), __ddiv_rn(L, l_max)),
                                 __dadd_rn(__dmul_rn(G.w_e, s_el), __dmul_rn(G.w_c, s_c)));
      score_max(T, s, c, G.q);
    }
}

__global__ void __launch_bounds__(kThreads) k_score_geometric_tri(Mesh M, Table T, Geo G, const int32_t* tris,
                                                                  int n_tris, const float* tri_dist,
                                                                  const float* tri_bary) {
  const int f = blockIdx.x * kThreads + threadIdx.x;
  if (f >= n_tris) return;
  const float d = tri_dist[f];
  if (!((double)d < G.d1)) return;
  const double prox = proximity(d, G.d1);
  const int32_t* tv = tris + 3 * f;
  double l_max = 1.0e-20;
  for (int k = 0; k < 3; ++k) l_max = fmax(l_max, edge_len(M.x, tv[k], tv[(k + 1) % 3]));
  for (int k = 0; k < 3; ++k) {
    // edge opposite corner k
    const int a = tv[(k + 1) % 3], b = tv[(k + 2) % 3];
    if (M.inv_mass[a] == 0.f && M.inv_mass[b] == 0.f) continue;
    const double L = edge_len(M.x, a, b);
    if (L < __dmul_rn(2.0, G.h_min)) continue;
    const int s = table_find(T.keys, T.mask, ekey(a, b));
    if (s < 0) continue;
    const double w = __dsub_rn(1.0, (double)tri_bary[3 * f + k]);
    const double c = __dmul_rn(__dmul_rn(__dmul_rn(__dmul_rn(__ddiv_rn(L, G.h_min), __ddiv_rn(L, l_max)), G.w_t), prox), w);
    score_max(T, s, c, G.q);
  }
}

__global__ void k_finalize_scores(Table T, long long pen_int) {
  const uint32_t s = blockIdx.x * kThreads + threadIdx.x;
  if (s > T.mask) return;
  if (T.keys[s] == 0ull) return;
  long long v = (long long)T.score[s];
  if (v > kScoreCap) v = kScoreCap;
  if (T.pen[s] && pen_int > v) v = pen_int;
  T.score[s] = (unsigned long long)v;
  T.state[s] = v > (long long)kScoreOne ? kAvailable : kUnavailable;
}

__device__ __forceinline__ bool better(long long s1, uint64_t k1, long long s2, uint64_t k2) {
  return s1 > s2 || (s1 == s2 && k1 > k2);
}

// every unsplit tet votes for its best available edge
__global__ void __launch_bounds__(kThreads) k_vote(Mesh M, Table T, const int32_t* split) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= M.n_tets || split[t] >= 0) return;
  const int32_t* tv = M.tets + 4 * t;
  int best = -1;
  long long bs = 0;
  uint64_t bk = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j) {
      const uint64_t k = ekey(tv[i], tv[j]);
      const int s = table_find(T.keys, T.mask, k);
      if (s < 0 || T.state[s] != kAvailable) continue;
      const long long sc = (long long)T.score[s];
      if (best < 0 || better(sc, k, bs, bk)) {
        best = s;
        bs = sc;
        bk = k;
      }
    }
  if (best >= 0) atomicAdd(&T.votes[best], 1);
}

// an available edge with the votes of all its tets is selected
__global__ void k_select(Table T, int32_t* changed) {
  const uint32_t s = blockIdx.x * kThreads + threadIdx.x;
  if (s > T.mask) return;
  if (T.keys[s] != 0ull && T.state[s] == kAvailable && T.votes[s] > 0 && T.votes[s] == T.cnt[s]) {
    T.state[s] = kSelected;
    atomicAdd(changed, 1);
  }
  T.votes[s] = 0;
}

// tets on a selected edge are split; their other edges become unavailable
__global__ void __launch_bounds__(kThreads) k_claim(Mesh M, Table T, int32_t* split) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= M.n_tets || split[t] >= 0) return;
  const int32_t* tv = M.tets + 4 * t;
  int sel = -1;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j) {
      const int s = table_find(T.keys, T.mask, ekey(tv[i], tv[j]));
      if (s >= 0 && T.state[s] == kSelected) sel = s;
    }
  if (sel < 0) return;
  split[t] = sel;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j) {
      const int s = table_find(T.keys, T.mask, ekey(tv[i], tv[j]));
      if (s >= 0 && s != sel && T.state[s] == kAvailable) T.state[s] = kUnavailable;
    }
}

// per tet: children (1 or 2) and whether it numbers a new vertex
__global__ void k_counts(Mesh M, Table T, const int32_t* split, int32_t* nchild, int32_t* nnew) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= M.n_tets) return;
  const int s = split[t];
  nchild[t] = s >= 0 ? 2 : 1;
  nnew[t] = (s >= 0 && T.owner[s] == t) ? 1 : 0;
}

// new vertices: id = n_verts + scan, x / v / rest interpolated from the
// lower-id endpoint (claim_new_vertices, refinement.py:697-730)
__global__ void k_new_vertices(Mesh M, Table T, const int32_t* split, const int32_t* nnew, const int32_t* newscan,
                               int n_verts, float split_t, float* x, float* v, float* rest) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= M.n_tets || !nnew[t]) return;
  const int s = split[t];
  const int id = n_verts + newscan[t];
  T.newvid[s] = id;
  const uint64_t k = T.keys[s];
  const int a = (int)(k >> 32), b = (int)(k & 0xffffffffu);
  float* arrs[3] = {x, v, rest};
  for (int q = 0; q < 3; ++q) {
    float* p = arrs[q];
    for (int c = 0; c < 3; ++c)
      p[3 * id + c] = __fadd_rn(p[3 * a + c], __fmul_rn(__fsub_rn(p[3 * b + c], p[3 * a + c]), split_t));
  }
}

// tets (scatter_tets, refinement.py:733-824): children replace corner i / j
// of the split edge (i > j in loop order); per-tet data copied, DmInv from
// the rest positions
__global__ void __launch_bounds__(kThreads) k_scatter_tets(Mesh M, Table T, const int32_t* split,
                                                           const int32_t* cscan, const float* rest,
                                                           const float* stretch, const float* lam,
                                                           const float* mat, const float* dminv, int32_t* o_tets,
                                                           float* o_stretch, float* o_lam, float* o_mat,
                                                           float* o_dminv) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= M.n_tets) return;
  const int32_t* tv = M.tets + 4 * t;
  const int base = cscan[t];
  const int s = split[t];
  int li = 0, lj = 0, nv = -1;
  if (s >= 0) {
    const uint64_t k = T.keys[s];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < i; ++j)
        if (ekey(tv[i], tv[j]) == k) {
          li = i;
          lj = j;
        }
    nv = T.newvid[s];
  }
  const int nc = s >= 0 ? 2 : 1;
  for (int c = 0; c < nc; ++c) {
    const int o = base + c;
    int w[4] = {tv[0], tv[1], tv[2], tv[3]};
    if (s >= 0) w[c == 0 ? li : lj] = nv;
    for (int q = 0; q < 4; ++q) o_tets[4 * o + q] = w[q];
    for (int q = 0; q < 6; ++q) {
      o_stretch[6 * o + q] = stretch[6 * t + q];
      o_lam[6 * o + q] = lam[6 * t + q];
    }
    for (int q = 0; q < 3; ++q) o_mat[3 * o + q] = mat[3 * t + q];
    if (s < 0) {
      for (int q = 0; q < 9; ++q) o_dminv[9 * o + q] = dminv[9 * t + q];
    } else {
      float Dm[9], Di[9];
      for (int r = 0; r < 3; ++r)
        for (int e = 0; e < 3; ++e) Dm[3 * r + e] = rest[3 * w[e + 1] + r] - rest[3 * w[0] + r];
      inv3(Dm, Di);
      for (int q = 0; q < 9; ++q) o_dminv[9 * o + q] = Di[q];
    }
  }
}

// surface triangles on a split edge (get_tri_candidate / scatter_tris)
__global__ void k_tri_counts(const int32_t* tris, int n_tris, Table T, int32_t* tsplit, int32_t* nchild) {
  const int f = blockIdx.x * kThreads + threadIdx.x;
  if (f >= n_tris) return;
  const int32_t* tv = tris + 3 * f;
  int sel = -1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < i; ++j) {
      const int s = table_find(T.keys, T.mask, ekey(tv[i], tv[j]));
      if (s >= 0 && T.state[s] == kSelected) sel = s;
    }
  tsplit[f] = sel;
  nchild[f] = sel >= 0 ? 2 : 1;
}

__global__ void k_scatter_tris(const int32_t* tris, int n_tris, Table T, const int32_t* tsplit, const int32_t* cscan,
                               int32_t* o_tris) {
  const int f = blockIdx.x * kThreads + threadIdx.x;
  if (f >= n_tris) return;
  const int32_t* tv = tris + 3 * f;
  const int s = tsplit[f];
  const int base = cscan[f];
  if (s < 0) {
    for (int q = 0; q < 3; ++q) o_tris[3 * base + q] = tv[q];
    return;
  }
  const uint64_t k = T.keys[s];
  int li = 0, lj = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < i; ++j)
      if (ekey(tv[i], tv[j]) == k) {
        li = i;
        lj = j;
      }
  for (int c = 0; c < 2; ++c)
    for (int q = 0; q < 3; ++q) o_tris[3 * (base + c) + q] = (q == (c == 0 ? li : lj)) ? T.newvid[s] : tv[q];
}

}  // namespace

extern "C" {
// scan helper from mfem_collide.hip
int mfem_scan_i32(const int32_t* in, int32_t* out, int n, int inclusive, int32_t* total, void* workspace,
                  size_t workspace_bytes, void* stream);
size_t mfem_scan_workspace_bytes(int n);
const char* mfem_collide_last_error(void);
}

namespace {
struct RLayout {
  uint32_t tsize;
  size_t off_keys, off_score, off_pen, off_cnt, off_votes, off_state, off_owner, off_newvid, off_split, off_nchild,
      off_nnew, off_cscan, off_newscan, off_tsplit, off_tchild, off_tscan, off_totals, off_dens, off_scan, bytes;
};
RLayout r_layout(long max_tets, long max_tris) {
  RLayout L{};
  uint32_t ts = 1024;
  while ((long)ts < 2 * 6 * max_tets && ts < (1u << 30)) ts <<= 1;
  L.tsize = ts;
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t o = 0;
  L.off_keys = o; o += al((size_t)ts * 8);
  L.off_score = o; o += al((size_t)ts * 8);
  L.off_pen = o; o += al((size_t)ts * 4);
  L.off_cnt = o; o += al((size_t)ts * 4);
  L.off_votes = o; o += al((size_t)ts * 4);
  L.off_state = o; o += al((size_t)ts * 4);
  L.off_owner = o; o += al((size_t)ts * 4);
  L.off_newvid = o; o += al((size_t)ts * 4);
  L.off_split = o; o += al((size_t)max_tets * 4);
  L.off_nchild = o; o += al((size_t)max_tets * 4);
  L.off_nnew = o; o += al((size_t)max_tets * 4);
  L.off_cscan = o; o += al((size_t)max_tets * 4);
  L.off_newscan = o; o += al((size_t)max_tets * 4);
  L.off_tsplit = o; o += al((size_t)(max_tris + 1) * 4);
  L.off_tchild = o; o += al((size_t)(max_tris + 1) * 4);
  L.off_tscan = o; o += al((size_t)(max_tris + 1) * 4);
  L.off_totals = o; o += al(64);
  L.off_dens = o; o += al(((size_t)max_tets / kThreads + 2) * 8);
  L.off_scan = o;
  const long m = max_tets > max_tris ? max_tets : max_tris;
  o += mfem_scan_workspace_bytes((int)m);
  L.bytes = o;
  return L;
}
}  // namespace

extern "C" {

const char* mfem_refine_last_error(void) { return g_err.c_str(); }

size_t mfem_refine_workspace_bytes(int max_tets, int max_tris) { return r_layout(max_tets, max_tris).bytes; }

int mfem_refine_select(const mfem_refine_args* A, void* workspace, size_t workspace_bytes, int32_t* status,
                       int32_t* changed, int phase, void* stream) {
  if (!A || !workspace || !status || !changed) return fail(MFEM_EINVAL, "null argument");
  if (A->n_tets <= 0 || A->n_tets > A->max_tets || A->n_tris < 0 || A->n_tris > A->max_tris)
    return fail(MFEM_EINVAL, "bad tet / triangle counts");
  if (!(A->threshold > 0.f)) return fail(MFEM_EINVAL, "threshold must be > 0");
  if (A->scoring != 0 && A->scoring != 1) return fail(MFEM_EINVAL, "scoring must be 0 (legacy) or 1 (geometric)");
  if (A->scoring == 1 && (!(A->min_edge_length > 0.f) || !(A->contact_d1 > 0.f)))
    return fail(MFEM_EINVAL, "geometric scoring needs min_edge_length > 0 and contact_d1 > 0");
  if (A->scoring == 1 && (!A->materials || !A->dminv || !A->tet_score))
    return fail(MFEM_EINVAL, "geometric scoring needs tet_score, materials and dminv");
  if (A->scoring == 1 && A->vertex_sid && !A->vertex_dist)
    return fail(MFEM_EINVAL, "vertex_sid needs vertex_dist");
  const RLayout L = r_layout(A->max_tets, A->max_tris);
  if (workspace_bytes < L.bytes) return fail(MFEM_ECAPACITY, "refinement workspace too small");
  char* ws = static_cast<char*>(workspace);
  Table T{reinterpret_cast<unsigned long long*>(ws + L.off_keys), reinterpret_cast<unsigned long long*>(ws + L.off_score),
          reinterpret_cast<int32_t*>(ws + L.off_pen), reinterpret_cast<int32_t*>(ws + L.off_cnt),
          reinterpret_cast<int32_t*>(ws + L.off_votes), reinterpret_cast<int32_t*>(ws + L.off_state),
          reinterpret_cast<int32_t*>(ws + L.off_owner), reinterpret_cast<int32_t*>(ws + L.off_newvid), L.tsize - 1};
  int32_t* split = reinterpret_cast<int32_t*>(ws + L.off_split);
  Mesh M{A->tets, A->n_tets, A->x, A->inv_mass, A->tet_score, A->vertex_score, A->shape_type, A->shape_xf,
         A->shape_scale, A->n_shapes};
  hipStream_t st = (hipStream_t)stream;
  if (phase == 0) {  // build the table and scores, clear the selection
    hipLaunchKernelGGL(k_fill64, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, T.keys, (long)L.tsize);
    hipLaunchKernelGGL(k_fill64, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, T.score, (long)L.tsize);
    hipLaunchKernelGGL(k_fill, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, T.pen, (long)L.tsize, 0);
    hipLaunchKernelGGL(k_fill, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, T.cnt, (long)L.tsize, 0);
    hipLaunchKernelGGL(k_fill, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, T.votes, (long)L.tsize, 0);
    hipLaunchKernelGGL(k_fill, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, T.owner, (long)L.tsize, 0x7fffffff);
    hipLaunchKernelGGL(k_fill, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, split, (long)A->n_tets, -1);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kThreads), 0, st, status, 1L, 0);
    if (int rc = launched()) return rc;
    const double q = kScoreOne / (double)A->threshold;
    if (A->scoring == 1) {
      double* part = reinterpret_cast<double*>(ws + L.off_dens);
      double* total = part + (A->max_tets / kThreads + 1);
      hipLaunchKernelGGL(k_density_partial, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, A->tet_score,
                         A->materials, A->dminv, A->n_tets, part);
      hipLaunchKernelGGL(k_density_total, dim3(1), dim3(kThreads), 0, st, (const double*)part, (int)blocks(A->n_tets),
                         total);
      Geo G{A->materials, A->dminv, A->vertex_dist, A->vertex_sid, total, (double)A->min_edge_length,
            (double)A->contact_d1, (double)A->w_elastic, (double)A->w_vertex_contact, (double)A->w_tri_contact, q};
      hipLaunchKernelGGL(k_score_geometric, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, M, T, G, status);
      if (A->tri_dist && A->tri_bary && A->n_tris > 0)
        hipLaunchKernelGGL(k_score_geometric_tri, dim3(blocks(A->n_tris)), dim3(kThreads), 0, st, M, T, G, A->tris,
                           A->n_tris, A->tri_dist, A->tri_bary);
    } else {
      hipLaunchKernelGGL(k_score, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, M, T, A->w_tet, A->w_vertex, q,
                         status);
    }
    if (int rc = launched()) return rc;
    long long pen_int = __builtin_llrint((double)A->penetrating_score * q);
    if (pen_int < 0) pen_int = 0;
    if (pen_int > kScoreCap) pen_int = kScoreCap;
    hipLaunchKernelGGL(k_finalize_scores, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, T, pen_int);
    return launched();
  }
  // one selection round: vote, select (changed += selections), claim
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(kThreads), 0, st, changed, 1L, 0);
  hipLaunchKernelGGL(k_vote, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, M, T, (const int32_t*)split);
  hipLaunchKernelGGL(k_select, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, T, changed);
  hipLaunchKernelGGL(k_claim, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, M, T, split);
  return launched();
}

int mfem_refine_count(const mfem_refine_args* A, void* workspace, size_t workspace_bytes, int32_t* totals,
                      void* stream) {
  if (!A || !workspace || !totals) return fail(MFEM_EINVAL, "null argument");
  const RLayout L = r_layout(A->max_tets, A->max_tris);
  if (workspace_bytes < L.bytes) return fail(MFEM_ECAPACITY, "refinement workspace too small");
  char* ws = static_cast<char*>(workspace);
  Table T{reinterpret_cast<unsigned long long*>(ws + L.off_keys), reinterpret_cast<unsigned long long*>(ws + L.off_score),
          reinterpret_cast<int32_t*>(ws + L.off_pen), reinterpret_cast<int32_t*>(ws + L.off_cnt),
          reinterpret_cast<int32_t*>(ws + L.off_votes), reinterpret_cast<int32_t*>(ws + L.off_state),
          reinterpret_cast<int32_t*>(ws + L.off_owner), reinterpret_cast<int32_t*>(ws + L.off_newvid), L.tsize - 1};
  Mesh M{A->tets, A->n_tets, A->x, A->inv_mass, A->tet_score, A->vertex_score, A->shape_type, A->shape_xf,
         A->shape_scale, A->n_shapes};
  int32_t* split = reinterpret_cast<int32_t*>(ws + L.off_split);
  int32_t* nchild = reinterpret_cast<int32_t*>(ws + L.off_nchild);
  int32_t* nnew = reinterpret_cast<int32_t*>(ws + L.off_nnew);
  int32_t* cscan = reinterpret_cast<int32_t*>(ws + L.off_cscan);
  int32_t* newscan = reinterpret_cast<int32_t*>(ws + L.off_newscan);
  int32_t* tsplit = reinterpret_cast<int32_t*>(ws + L.off_tsplit);
  int32_t* tchild = reinterpret_cast<int32_t*>(ws + L.off_tchild);
  int32_t* tscan = reinterpret_cast<int32_t*>(ws + L.off_tscan);
  char* sws = ws + L.off_scan;
  const size_t sbytes = L.bytes - L.off_scan;
  hipStream_t st = (hipStream_t)stream;
  // totals: [0] tets after the pass, [1] new vertices, [2] triangles after the pass
  hipLaunchKernelGGL(k_counts, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, M, T, (const int32_t*)split, nchild,
                     nnew);
  if (int rc = launched()) return rc;
  if (mfem_scan_i32(nchild, cscan, A->n_tets, 0, totals, sws, sbytes, stream))
    return fail(MFEM_EHIP, mfem_collide_last_error());
  if (mfem_scan_i32(nnew, newscan, A->n_tets, 0, totals + 1, sws, sbytes, stream))
    return fail(MFEM_EHIP, mfem_collide_last_error());
  if (A->n_tris > 0) {
    hipLaunchKernelGGL(k_tri_counts, dim3(blocks(A->n_tris)), dim3(kThreads), 0, st, A->tris, A->n_tris, T, tsplit,
                       tchild);
    if (int rc = launched()) return rc;
    if (mfem_scan_i32(tchild, tscan, A->n_tris, 0, totals + 2, sws, sbytes, stream))
      return fail(MFEM_EHIP, mfem_collide_last_error());
  } else {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kThreads), 0, st, totals + 2, 1L, 0);
  }
  return launched();
}

int mfem_refine_apply(const mfem_refine_args* A, void* workspace, size_t workspace_bytes, const mfem_refine_out* O,
                      void* stream) {
  if (!A || !workspace || !O) return fail(MFEM_EINVAL, "null argument");
  const RLayout L = r_layout(A->max_tets, A->max_tris);
  if (workspace_bytes < L.bytes) return fail(MFEM_ECAPACITY, "refinement workspace too small");
  char* ws = static_cast<char*>(workspace);
  Table T{reinterpret_cast<unsigned long long*>(ws + L.off_keys), reinterpret_cast<unsigned long long*>(ws + L.off_score),
          reinterpret_cast<int32_t*>(ws + L.off_pen), reinterpret_cast<int32_t*>(ws + L.off_cnt),
          reinterpret_cast<int32_t*>(ws + L.off_votes), reinterpret_cast<int32_t*>(ws + L.off_state),
          reinterpret_cast<int32_t*>(ws + L.off_owner), reinterpret_cast<int32_t*>(ws + L.off_newvid), L.tsize - 1};
  Mesh M{A->tets, A->n_tets, A->x, A->inv_mass, A->tet_score, A->vertex_score, A->shape_type, A->shape_xf,
         A->shape_scale, A->n_shapes};
  const int32_t* split = reinterpret_cast<int32_t*>(ws + L.off_split);
  const int32_t* nnew = reinterpret_cast<int32_t*>(ws + L.off_nnew);
  const int32_t* cscan = reinterpret_cast<int32_t*>(ws + L.off_cscan);
  const int32_t* newscan = reinterpret_cast<int32_t*>(ws + L.off_newscan);
  const int32_t* tsplit = reinterpret_cast<int32_t*>(ws + L.off_tsplit);
  const int32_t* tscan = reinterpret_cast<int32_t*>(ws + L.off_tscan);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_new_vertices, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, M, T, split, nnew, newscan,
                     A->n_verts, A->split_t, O->x, O->v, O->rest);
  if (int rc = launched()) return rc;
  hipLaunchKernelGGL(k_scatter_tets, dim3(blocks(A->n_tets)), dim3(kThreads), 0, st, M, T, split, cscan,
                     (const float*)O->rest, A->stretch, A->lam, A->materials, A->dminv, O->tets, O->stretch, O->lam,
                     O->materials, O->dminv);
  if (int rc = launched()) return rc;
  if (A->n_tris > 0) {
    hipLaunchKernelGGL(k_scatter_tris, dim3(blocks(A->n_tris)), dim3(kThreads), 0, st, A->tris, A->n_tris, T, tsplit,
                       tscan, O->tris);
  }
  return launched();
}

}  // extern "C"
