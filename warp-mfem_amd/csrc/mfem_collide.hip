// mfem_collide.hip -- self-contact building blocks (gfx950): int32 prefix
// scan, spatial-hash broad phase for point-triangle and edge-edge
// candidates, the reference's soft-rigid 15-feature distance kernel, and the
// IPC barrier on candidate pairs.
//
// Determinism: the hash table is filled with atomics (bucket order varies),
// so every query writes its candidates into a segment sized by a counting
// pass and a prefix scan, then sorts the segment by id: the output lists
// are identical run to run and equal the brute-force sets in sorted order.
// A candidate pair is reported only in the cell that holds the minimum
// corner of the intersection of the two boxes, so each pair appears once.
#include <hip/hip_runtime.h>

#include <mutex>

#include <cstdint>
#include <string>

#include "../../include/mfem_amd.h"

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;  // per thread
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int kThreads = 256;

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
inline unsigned blocks(long n, int t = kThreads) { return (unsigned)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// int32 prefix scan: tile sums -> scan of the sums -> tile scans + offsets
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// exclusive scan of one value per thread over the block; returns the block total
__device__ int block_excl_scan(int v, int& total) {
  __shared__ int wsum[kScanThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / 64; ++k) {
    if (k < w) off += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  total = tot;
  return off + inc - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tile_sums(const int32_t* in, int n, int32_t* sums) {
  const long base = (long)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  int total;
  block_excl_scan(s, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const int32_t* in, int n, int32_t* out,
                                                             const int32_t* offsets, int inclusive,
                                                             int32_t* total_out) {
  const long base = (long)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems], s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int run = block_excl_scan(s, total) + (offsets ? offsets[blockIdx.x] : 0);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = inclusive ? run + v[k] : run;
    run += v[k];
  }
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) *total_out = run;
}

constexpr size_t scan_ws_bytes(long n) {
  size_t b = 0;
  long m = n;
  while (m > kScanTile) {
    m = (m + kScanTile - 1) / kScanTile;
    b += ((size_t)m * sizeof(int32_t) + 255) / 256 * 256;
  }
  return b;
}

// exclusive (or inclusive) scan of in[0..n) into out; total (device, may be
// null) receives the sum.  Recursion over tile sums, workspace from ws.
int scan_i32(const int32_t* in, int32_t* out, long n, int inclusive, int32_t* total, char* ws, size_t ws_bytes,
             hipStream_t st) {
  if (n <= 0) return MFEM_OK;
  const long tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanThreads), 0, st, in, (int)n, out, (const int32_t*)nullptr,
                       inclusive, total);
    return launched();
  }
  const size_t need = ((size_t)tiles * sizeof(int32_t) + 255) / 256 * 256;
  if (ws_bytes < need) return fail(MFEM_ECAPACITY, "scan workspace too small");
  int32_t* sums = reinterpret_cast<int32_t*>(ws);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(tiles), dim3(kScanThreads), 0, st, in, (int)n, sums);
  if (int rc = launched()) return rc;
  // exclusive scan of the tile sums in place (total = overall sum)
  if (int rc = scan_i32(sums, sums, tiles, 0, total, ws + need, ws_bytes - need, st)) return rc;
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, st, in, (int)n, out, sums, inclusive,
                     (int32_t*)nullptr);
  return launched();
}

// ---------------------------------------------------------------------------
// scratch of the allocation-free scan (mfem_scan_i32_noalloc): the tile sums
// of every recursion level for inputs of up to MFEM_SCAN_NOALLOC_MAX values,
// 65536 sums at the first level and 64 at the second (256-byte aligned
// levels, as scan_i32 lays them out).  One copy per GPU, held until the
// process ends; every call on a GPU uses that GPU's copy, so calls that
// overlap on one GPU (two streams at once) race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 262144 + 256;
static_assert(MFEM_SCAN_NOALLOC_MAX / kScanTile == 65536, "first level of tile sums");
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's scratch (the array's host-side name is no
// device address): looked up once per device, with that device current
int noalloc_scratch(char** ws) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(MFEM_EHIP, "mfem_scan_i32_noalloc: no current device");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(MFEM_EHIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
    g_scratch_ptr[dev] = ptr;
  }
  *ws = static_cast<char*>(g_scratch_ptr[dev]);
  return MFEM_OK;
}
static_assert(scan_ws_bytes(MFEM_SCAN_NOALLOC_MAX) == kNoAllocScratchBytes, "scratch = the scan's tile sums at the maximum");

// ---------------------------------------------------------------------------
// geometry (src/mfem/collision.py:220-322)
// ---------------------------------------------------------------------------
struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 ld3(const float* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 mul(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len(f3 a) { return sqrtf(dot(a, a)); }

// closest point of triangle abc to x with its barycentric weights
__device__ f3 closest_on_triangle(f3 x, f3 a, f3 b, f3 c, float* w) {
  const f3 ab = sub(b, a), ac = sub(c, a), ax = sub(x, a);
  const float d1 = dot(ab, ax), d2 = dot(ac, ax);
  if (d1 <= 0.f && d2 <= 0.f) { w[0] = 1.f; w[1] = 0.f; w[2] = 0.f; return a; }
  const f3 bx = sub(x, b);
  const float d3 = dot(ab, bx), d4 = dot(ac, bx);
  if (d3 >= 0.f && d4 <= d3) { w[0] = 0.f; w[1] = 1.f; w[2] = 0.f; return b; }
  const f3 cx = sub(x, c);
  const float d5 = dot(ab, cx), d6 = dot(ac, cx);
  if (d6 >= 0.f && d5 <= d6) { w[0] = 0.f; w[1] = 0.f; w[2] = 1.f; return c; }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {
    const float v = d1 / (d1 - d3);
    w[0] = 1.f - v; w[1] = v; w[2] = 0.f;
    return add(a, mul(v, ab));
  }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {
    const float v = d2 / (d2 - d6);
    w[0] = 1.f - v; w[1] = 0.f; w[2] = v;
    return add(a, mul(v, ac));
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
    const float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    w[0] = 0.f; w[1] = 1.f - v; w[2] = v;
    return add(b, mul(v, sub(c, b)));
  }
  const float den = 1.f / (va + vb + vc);
  const float v = vb * den, u = vc * den;
  w[0] = 1.f - v - u; w[1] = v; w[2] = u;
  return add(a, add(mul(v, ab), mul(u, ac)));
}

// closest points of segments p1q1 / p2q2; s, t their parameters
__device__ void closest_seg_seg(f3 p1, f3 q1, f3 p2, f3 q2, f3& c1, f3& c2, float& s, float& t) {
  const f3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
  const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const float eps = 1.0e-8f;
  s = 0.f;
  t = 0.f;
  if (a <= eps && e <= eps) {
  } else if (a <= eps) {
    t = fminf(fmaxf(f / e, 0.f), 1.f);
  } else {
    const float c = dot(d1, r);
    if (e <= eps) {
      s = fminf(fmaxf(-c / a, 0.f), 1.f);
    } else {
      const float b = dot(d1, d2), den = a * e - b * b;
      s = den > eps ? fminf(fmaxf((b * f - c * e) / den, 0.f), 1.f) : 0.f;
      t = (b * s + f) / e;
      if (t < 0.f) {
        t = 0.f;
        s = fminf(fmaxf(-c / a, 0.f), 1.f);
      } else if (t > 1.f) {
        t = 1.f;
        s = fminf(fmaxf((b - c) / a, 0.f), 1.f);
      }
    }
  }
  c1 = add(p1, mul(s, d1));
  c2 = add(p2, mul(t, d2));
}

// atomic minimum of a non-negative float (its bit pattern orders like an int)
__device__ __forceinline__ void atomic_min_pos(float* p, float v) {
  atomicMin(reinterpret_cast<int*>(p), __float_as_int(v));
}

// the reference's rigid/soft candidate-pair distance kernel
// (src/mfem/collision.py:380-476): per soft particle the minimum of the 15
// feature distances below d_hat
__global__ void __launch_bounds__(kThreads) k_soft_rigid_distance(const float* points, const int32_t* tris,
                                                                  int rigid_points_count, int rigid_tri_count,
                                                                  const int32_t* pairs, const int32_t* pair_count,
                                                                  int max_pairs, float d_hat, float* distance) {
  const int tid = blockIdx.x * kThreads + threadIdx.x;
  const int n = min(pair_count[0], max_pairs);
  if (tid >= n) return;
  const int ta = pairs[2 * tid], tb = pairs[2 * tid + 1];
  const bool ar = ta < rigid_tri_count, br = tb < rigid_tri_count;
  if (ar == br) return;
  const int tr = ar ? ta : tb, ts = ar ? tb : ta;
  f3 r[3], sp[3];
  int p[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r[k] = ld3(points, tris[3 * tr + k]);
    const int si = tris[3 * ts + k];
    sp[k] = ld3(points, si);
    p[k] = si - rigid_points_count;
  }
  float w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {  // soft points vs the rigid triangle
    const float d = len(sub(sp[k], closest_on_triangle(sp[k], r[0], r[1], r[2], w)));
    if (d < d_hat) atomic_min_pos(distance + p[k], d);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {  // rigid points vs the soft triangle: all three soft vertices
    const float d = len(sub(r[k], closest_on_triangle(r[k], sp[0], sp[1], sp[2], w)));
    if (d < d_hat)
#pragma unroll
      for (int q = 0; q < 3; ++q) atomic_min_pos(distance + p[q], d);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {  // soft edges vs rigid edges
      f3 c1, c2;
      float s, t;
      closest_seg_seg(sp[i], sp[(i + 1) % 3], r[j], r[(j + 1) % 3], c1, c2, s, t);
      const float d = len(sub(c1, c2));
      if (d < d_hat) {
        atomic_min_pos(distance + p[i], d);
        atomic_min_pos(distance + p[(i + 1) % 3], d);
      }
    }
}

// ---------------------------------------------------------------------------
// spatial hash (uniform grid, Teschner et al. 2003 hashing of cell coords)
// ---------------------------------------------------------------------------
struct Box {
  f3 lo, hi;
};
struct Entry {
  int32_t id;
  int32_t pad;
  uint64_t key;
};
// cell coordinate, clamped to the 21-bit key range (far-away primitives
// share boundary cells: more candidates to test, none lost)
__device__ __forceinline__ int cell_of(float v, float inv_h) {
  return (int)fminf(fmaxf(floorf(v * inv_h), -1048575.f), 1048575.f);
}
__device__ __forceinline__ uint64_t cell_key(int ix, int iy, int iz) {
  return ((uint64_t)(ix & 0x1FFFFF) << 42) | ((uint64_t)(iy & 0x1FFFFF) << 21) | (uint64_t)(iz & 0x1FFFFF);
}
__device__ __forceinline__ uint32_t cell_hash(int ix, int iy, int iz, uint32_t mask) {
  return ((uint32_t)ix * 73856093u ^ (uint32_t)iy * 19349663u ^ (uint32_t)iz * 83492791u) & mask;
}
// box of nv vertices (ids) inflated by r, in the same FP32 operations as
// the brute-force reference (min / max, then -r / +r); with dx (FP64
// displacements, may be null) the box of the swept primitive: every vertex
// at x and at x + dx (so that the box holds the primitive at x + a dx for
// every a in [0, 1])
template <int NV>
__device__ Box prim_box(const float* x, const double* dx, const int32_t* ids, float r) {
  f3 lo = ld3(x, ids[0]), hi = lo;
#pragma unroll
  for (int k = 1; k < NV; ++k) {
    const f3 v = ld3(x, ids[k]);
    lo = {fminf(lo.x, v.x), fminf(lo.y, v.y), fminf(lo.z, v.z)};
    hi = {fmaxf(hi.x, v.x), fmaxf(hi.y, v.y), fmaxf(hi.z, v.z)};
  }
  if (dx) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const f3 v = ld3(x, ids[k]);
      const double* d = dx + 3 * (size_t)ids[k];
      const f3 w = {v.x + (float)d[0], v.y + (float)d[1], v.z + (float)d[2]};
      lo = {fminf(lo.x, w.x), fminf(lo.y, w.y), fminf(lo.z, w.z)};
      hi = {fmaxf(hi.x, w.x), fmaxf(hi.y, w.y), fmaxf(hi.z, w.z)};
    }
  }
  return {{lo.x - r, lo.y - r, lo.z - r}, {hi.x + r, hi.y + r, hi.z + r}};
}
__device__ __forceinline__ bool overlap(const Box& a, const Box& b) {
  return a.lo.x <= b.hi.x && b.lo.x <= a.hi.x && a.lo.y <= b.hi.y && b.lo.y <= a.hi.y && a.lo.z <= b.hi.z &&
         b.lo.z <= a.hi.z;
}

struct Grid {
  float inv_h;
  uint32_t mask;  // table size - 1 (power of two)
  int max_span;   // cells per axis a primitive may span
};

__device__ __forceinline__ bool cell_range(const Box& b, const Grid& g, int* lo, int* hi) {
  lo[0] = cell_of(b.lo.x, g.inv_h); hi[0] = cell_of(b.hi.x, g.inv_h);
  lo[1] = cell_of(b.lo.y, g.inv_h); hi[1] = cell_of(b.hi.y, g.inv_h);
  lo[2] = cell_of(b.lo.z, g.inv_h); hi[2] = cell_of(b.hi.z, g.inv_h);
  return hi[0] - lo[0] < g.max_span && hi[1] - lo[1] < g.max_span && hi[2] - lo[2] < g.max_span;
}

__global__ void k_fill_i32(int32_t* p, long n, int32_t v) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

// insertion pass: MODE 0 counts per bucket, MODE 1 writes entries
template <int NV, int MODE>
__global__ void __launch_bounds__(kThreads) k_insert(const float* x, const double* dx, const int32_t* prims, int n,
                                                     float r, Grid g,
                                                     int32_t* counts, const int32_t* start, int32_t* cursor,
                                                     Entry* entries, int cap, int32_t* status) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const Box b = prim_box<NV>(x, dx, prims + NV * i, r);
  int lo[3], hi[3];
  if (!cell_range(b, g, lo, hi)) {
    atomicOr(status, MFEM_BP_SPAN);
    return;
  }
  for (int ix = lo[0]; ix <= hi[0]; ++ix)
    for (int iy = lo[1]; iy <= hi[1]; ++iy)
      for (int iz = lo[2]; iz <= hi[2]; ++iz) {
        const uint32_t h = cell_hash(ix, iy, iz, g.mask);
        if (MODE == 0) {
          atomicAdd(&counts[h], 1);
        } else {
          const int slot = start[h] + atomicAdd(&cursor[h], 1);
          if (slot < cap) entries[slot] = {i, 0, cell_key(ix, iy, iz)};
          else atomicOr(status, MFEM_BP_ENTRIES);
        }
      }
}

// query pass for point-triangle (QNV = 1, ENV = 3) or edge-edge (QNV = 2,
// ENV = 2, only partners with a larger id); MODE 0 counts, MODE 1 writes and
// sorts the query's segment
template <int QNV, int ENV, int MODE>
__global__ void __launch_bounds__(kThreads) k_query(const float* x, const double* dx, const int32_t* qprims, int nq,
                                                    const int32_t* eprims, float r, Grid g, const int32_t* counts,
                                                    const int32_t* start, const Entry* entries, int cap,
                                                    int32_t* qcount, const int32_t* qstart, int32_t* pairs,
                                                    int max_pairs, int32_t* status) {
  const int q = blockIdx.x * kThreads + threadIdx.x;
  if (q >= nq) return;
  const int32_t* qv = qprims + QNV * q;
  const Box bq = prim_box<QNV>(x, dx, qv, r);
  int lo[3], hi[3];
  if (!cell_range(bq, g, lo, hi)) {
    atomicOr(status, MFEM_BP_SPAN);
    if (MODE == 0) qcount[q] = 0;
    return;
  }
  const int qid = QNV == 1 ? qv[0] : q;  // point-triangle pairs carry the vertex id
  int found = 0;
  const int base = MODE == 1 ? qstart[q] : 0;
  for (int ix = lo[0]; ix <= hi[0]; ++ix)
    for (int iy = lo[1]; iy <= hi[1]; ++iy)
      for (int iz = lo[2]; iz <= hi[2]; ++iz) {
        const uint32_t h = cell_hash(ix, iy, iz, g.mask);
        const uint64_t key = cell_key(ix, iy, iz);
        const int j0 = start[h], j1 = min(start[h] + counts[h], cap);
        for (int j = j0; j < j1; ++j) {
          const Entry en = entries[j];
          if (en.key != key) continue;
          const int e = en.id;
          if (ENV == 2 && QNV == 2 && e <= q) continue;
          const int32_t* ev = eprims + ENV * e;
          bool shared = false;
#pragma unroll
          for (int a = 0; a < QNV; ++a)
#pragma unroll
            for (int b = 0; b < ENV; ++b) shared |= qv[a] == ev[b];
          if (shared) continue;
          const Box be = prim_box<ENV>(x, dx, ev, r);
          if (!overlap(bq, be)) continue;
          // report in the cell holding the minimum corner of the intersection
          if (cell_of(fmaxf(bq.lo.x, be.lo.x), g.inv_h) != ix || cell_of(fmaxf(bq.lo.y, be.lo.y), g.inv_h) != iy ||
              cell_of(fmaxf(bq.lo.z, be.lo.z), g.inv_h) != iz)
            continue;
          if (MODE == 1) {
            const int k = base + found;
            if (k < max_pairs) {
              pairs[2 * k] = qid;
              pairs[2 * k + 1] = e;
            }
          }
          ++found;
        }
      }
  if (MODE == 0) {
    qcount[q] = found;
    return;
  }
  // sort this query's segment by partner id (insertion sort; segments are short)
  const int n = min(found, max_pairs - base);
  if (base + found > max_pairs) atomicOr(status, MFEM_BP_PAIRS);
  for (int a = 1; a < n; ++a) {
    const int v = pairs[2 * (base + a) + 1];
    int b = a - 1;
    while (b >= 0 && pairs[2 * (base + b) + 1] > v) {
      pairs[2 * (base + b + 1) + 1] = pairs[2 * (base + b) + 1];
      --b;
    }
    pairs[2 * (base + b + 1) + 1] = v;
  }
}

// ---------------------------------------------------------------------------
// IPC barrier on candidate pairs: energy and gradient per pair
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ipc_barrier(float d, float dh, float& e, float& de) {
  const float l = logf(d / dh), u = d - dh;
  e = -u * u * l;
  de = -2.f * u * l - u * u / d;
}

// point-triangle (KIND 0, pairs = (vertex, tri)) or edge-edge (KIND 1,
// pairs = (edge, edge)): distance, kappa * b(d), and the gradient of that
// energy with respect to the pair's four vertices (ids out too)
template <int KIND>
__global__ void __launch_bounds__(kThreads) k_pair_barrier(const float* x, const int32_t* pairs,
                                                           const int32_t* n_pairs, int max_pairs, const int32_t* tris,
                                                           const int32_t* edges, float kappa, float d_hat,
                                                           float* dist, double* energy, float* grad, int32_t* vids) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= min(n_pairs[0], max_pairs)) return;
  int v[4];
  float cw[4];  // dd/dx_v = cw[v] * n
  f3 n;
  float d;
  if (KIND == 0) {
    v[0] = pairs[2 * k];
    const int t = pairs[2 * k + 1];
    v[1] = tris[3 * t]; v[2] = tris[3 * t + 1]; v[3] = tris[3 * t + 2];
    float w[3];
    const f3 p = ld3(x, v[0]);
    const f3 c = closest_on_triangle(p, ld3(x, v[1]), ld3(x, v[2]), ld3(x, v[3]), w);
    const f3 dv = sub(p, c);
    d = len(dv);
    n = d > 0.f ? mul(1.f / d, dv) : f3{0.f, 0.f, 0.f};
    cw[0] = 1.f; cw[1] = -w[0]; cw[2] = -w[1]; cw[3] = -w[2];
  } else {
    const int e1 = pairs[2 * k], e2 = pairs[2 * k + 1];
    v[0] = edges[2 * e1]; v[1] = edges[2 * e1 + 1]; v[2] = edges[2 * e2]; v[3] = edges[2 * e2 + 1];
    f3 c1, c2;
    float s, t;
    closest_seg_seg(ld3(x, v[0]), ld3(x, v[1]), ld3(x, v[2]), ld3(x, v[3]), c1, c2, s, t);
    const f3 dv = sub(c1, c2);
    d = len(dv);
    n = d > 0.f ? mul(1.f / d, dv) : f3{0.f, 0.f, 0.f};
    cw[0] = 1.f - s; cw[1] = s; cw[2] = -(1.f - t); cw[3] = -t;
  }
  float e = 0.f, de = 0.f;
  if (d > 0.f && d < d_hat) ipc_barrier(d, d_hat, e, de);
  dist[k] = d;
  energy[k] = (double)kappa * (double)e;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    vids[4 * k + a] = v[a];
    grad[12 * k + 3 * a] = kappa * de * cw[a] * n.x;
    grad[12 * k + 3 * a + 1] = kappa * de * cw[a] * n.y;
    grad[12 * k + 3 * a + 2] = kappa * de * cw[a] * n.z;
  }
}

uint32_t table_size(long inserts) {
  uint32_t s = 1024;
  while ((long)s < 2 * inserts && s < (1u << 30)) s <<= 1;
  return s;
}

struct BpLayout {
  uint32_t tsize;
  size_t off_counts, off_start, off_cursor, off_entries, off_qcount, off_qstart, off_total, off_scan, bytes;
  int cap;
};
BpLayout bp_layout(long n_insert, long n_query, int max_span) {
  BpLayout L{};
  const long per = (long)max_span * max_span * max_span;
  L.cap = (int)(n_insert * per < (1L << 30) ? n_insert * per : (1L << 30));
  L.tsize = table_size(n_insert * 8);
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t o = 0;
  L.off_counts = o; o += al((size_t)L.tsize * 4);
  L.off_start = o; o += al((size_t)L.tsize * 4);
  L.off_cursor = o; o += al((size_t)L.tsize * 4);
  L.off_entries = o; o += al((size_t)L.cap * sizeof(Entry));
  L.off_qcount = o; o += al((size_t)(n_query + 1) * 4);
  L.off_qstart = o; o += al((size_t)(n_query + 1) * 4);
  L.off_total = o; o += al(16);
  L.off_scan = o;
  const long m = (long)L.tsize > n_query ? (long)L.tsize : n_query;
  o += scan_ws_bytes(m) + 256;
  L.bytes = o;
  return L;
}

template <int QNV, int ENV>
int broadphase(const float* x, const double* dx, const int32_t* qprims, int nq, const int32_t* eprims, int ne,
               float d_hat, float cell,
               int max_span, int32_t* pairs, int max_pairs, int32_t* n_pairs, int32_t* status, void* workspace,
               size_t ws_bytes, hipStream_t st) {
  if (!x || !qprims || !eprims || !pairs || !n_pairs || !status || !workspace) return fail(MFEM_EINVAL, "null pointer");
  if (nq < 0 || ne < 0 || max_pairs < 0 || !(d_hat >= 0.f) || !(cell > 0.f) || max_span < 1 || max_span > 8)
    return fail(MFEM_EINVAL, "bad size, d_hat, cell or max_span");
  const BpLayout L = bp_layout(ne, nq, max_span);
  if (ws_bytes < L.bytes) return fail(MFEM_ECAPACITY, "broad-phase workspace too small");
  char* ws = static_cast<char*>(workspace);
  int32_t* counts = reinterpret_cast<int32_t*>(ws + L.off_counts);
  int32_t* start = reinterpret_cast<int32_t*>(ws + L.off_start);
  int32_t* cursor = reinterpret_cast<int32_t*>(ws + L.off_cursor);
  Entry* entries = reinterpret_cast<Entry*>(ws + L.off_entries);
  int32_t* qcount = reinterpret_cast<int32_t*>(ws + L.off_qcount);
  int32_t* qstart = reinterpret_cast<int32_t*>(ws + L.off_qstart);
  int32_t* total = reinterpret_cast<int32_t*>(ws + L.off_total);
  char* sws = ws + L.off_scan;
  const size_t sws_bytes = L.bytes - L.off_scan;
  const Grid g{1.f / cell, L.tsize - 1, max_span};
  const float r = 0.5f * d_hat;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, counts, (long)L.tsize, 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, cursor, (long)L.tsize, 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kThreads), 0, st, status, 1L, 0);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kThreads), 0, st, n_pairs, 1L, 0);
  if (int rc = launched()) return rc;
  if (ne > 0) {
    hipLaunchKernelGGL((k_insert<ENV, 0>), dim3(blocks(ne)), dim3(kThreads), 0, st, x, dx, eprims, ne, r, g, counts,
                       (const int32_t*)nullptr, cursor, entries, L.cap, status);
    if (int rc = launched()) return rc;
    if (int rc = scan_i32(counts, start, L.tsize, 0, total, sws, sws_bytes, st)) return rc;
    hipLaunchKernelGGL((k_insert<ENV, 1>), dim3(blocks(ne)), dim3(kThreads), 0, st, x, dx, eprims, ne, r, g, counts,
                       start, cursor, entries, L.cap, status);
    if (int rc = launched()) return rc;
  } else {
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks(L.tsize)), dim3(kThreads), 0, st, start, (long)L.tsize, 0);
  }
  if (nq == 0) return launched();
  hipLaunchKernelGGL((k_query<QNV, ENV, 0>), dim3(blocks(nq)), dim3(kThreads), 0, st, x, dx, qprims, nq, eprims, r, g,
                     counts, start, entries, L.cap, qcount, (const int32_t*)nullptr, pairs, max_pairs, status);
  if (int rc = launched()) return rc;
  if (int rc = scan_i32(qcount, qstart, nq, 0, n_pairs, sws, sws_bytes, st)) return rc;
  hipLaunchKernelGGL((k_query<QNV, ENV, 1>), dim3(blocks(nq)), dim3(kThreads), 0, st, x, dx, qprims, nq, eprims, r, g,
                     counts, start, entries, L.cap, qcount, qstart, pairs, max_pairs, status);
  return launched();
}

}  // namespace

extern "C" {

const char* mfem_collide_last_error(void) { return g_err.c_str(); }

size_t mfem_scan_workspace_bytes(int n) { return scan_ws_bytes(n) + 256; }

int mfem_scan_i32_noalloc(const int32_t* in, int32_t* out, int n, int inclusive, int32_t* total, void* stream) {
  if (n < 0 || (n > 0 && (!in || !out))) return fail(MFEM_EINVAL, "bad scan arguments");
  if (n > MFEM_SCAN_NOALLOC_MAX)
    return fail(MFEM_ECAPACITY, ("mfem_scan_i32_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                                 std::to_string(MFEM_SCAN_NOALLOC_MAX) + " values")
                                    .c_str());
  if (n == 0) {
    if (total) hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kThreads), 0, (hipStream_t)stream, total, 1L, 0);
    return launched();
  }
  char* ws = nullptr;
  if (int rc = noalloc_scratch(&ws)) return rc;
  // every tile sum the scan reads is written by the same call first: nothing
  // in the scratch needs clearing
  return scan_i32(in, out, n, inclusive, total, ws, kNoAllocScratchBytes, (hipStream_t)stream);
}

int mfem_scan_i32(const int32_t* in, int32_t* out, int n, int inclusive, int32_t* total, void* workspace,
                  size_t workspace_bytes, void* stream) {
  if (n < 0 || (n > 0 && (!in || !out))) return fail(MFEM_EINVAL, "bad scan arguments");
  if (n == 0) {
    if (total) hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kThreads), 0, (hipStream_t)stream, total, 1L, 0);
    return launched();
  }
  return scan_i32(in, out, n, inclusive, total, static_cast<char*>(workspace), workspace_bytes, (hipStream_t)stream);
}

size_t mfem_broadphase_workspace_bytes(int n_insert, int n_query, int max_span) {
  return bp_layout(n_insert, n_query, max_span < 1 ? 1 : (max_span > 8 ? 8 : max_span)).bytes;
}

int mfem_broadphase_pt(const float* x, const int32_t* points, int n_points, const int32_t* tris, int n_tris,
                       float d_hat, float cell, int max_span, int32_t* pairs, int max_pairs, int32_t* n_pairs,
                       int32_t* status, void* workspace, size_t workspace_bytes, void* stream) {
  return broadphase<1, 3>(x, nullptr, points, n_points, tris, n_tris, d_hat, cell, max_span, pairs, max_pairs, n_pairs,
                          status, workspace, workspace_bytes, (hipStream_t)stream);
}

int mfem_broadphase_ee(const float* x, const int32_t* edges, int n_edges, float d_hat, float cell, int max_span,
                       int32_t* pairs, int max_pairs, int32_t* n_pairs, int32_t* status, void* workspace,
                       size_t workspace_bytes, void* stream) {
  return broadphase<2, 2>(x, nullptr, edges, n_edges, edges, n_edges, d_hat, cell, max_span, pairs, max_pairs, n_pairs,
                          status, workspace, workspace_bytes, (hipStream_t)stream);
}

int mfem_broadphase_swept(int kind, const float* x, const double* dx, const int32_t* prims, int n_prims,
                          const int32_t* tris, int n_tris, float d_hat, float cell, int max_span, int32_t* pairs,
                          int max_pairs, int32_t* n_pairs, int32_t* status, void* workspace, size_t workspace_bytes,
                          void* stream) {
  if (kind == 0)
    return broadphase<1, 3>(x, dx, prims, n_prims, tris, n_tris, d_hat, cell, max_span, pairs, max_pairs, n_pairs,
                            status, workspace, workspace_bytes, (hipStream_t)stream);
  if (kind == 1)
    return broadphase<2, 2>(x, dx, prims, n_prims, prims, n_prims, d_hat, cell, max_span, pairs, max_pairs, n_pairs,
                            status, workspace, workspace_bytes, (hipStream_t)stream);
  return fail(MFEM_EINVAL, "mfem_broadphase_swept: kind must be 0 (point-triangle) or 1 (edge-edge)");
}

int mfem_soft_rigid_distance(const float* points, const int32_t* tris, int rigid_points_count, int rigid_tri_count,
                             const int32_t* pairs, const int32_t* pair_count, int max_pairs, float d_hat,
                             float* distance, void* stream) {
  if (!points || !tris || !pairs || !pair_count || !distance || max_pairs < 0) return fail(MFEM_EINVAL, "bad argument");
  if (max_pairs == 0) return MFEM_OK;
  hipLaunchKernelGGL(k_soft_rigid_distance, dim3(blocks(max_pairs)), dim3(kThreads), 0, (hipStream_t)stream, points,
                     tris, rigid_points_count, rigid_tri_count, pairs, pair_count, max_pairs, d_hat, distance);
  return launched();
}

int mfem_pair_barrier(int kind, const float* x, const int32_t* pairs, const int32_t* n_pairs, int max_pairs,
                      const int32_t* tris, const int32_t* edges, float kappa, float d_hat, float* dist,
                      double* energy, float* grad, int32_t* vids, void* stream) {
  if ((kind != 0 && kind != 1) || !x || !pairs || !n_pairs || max_pairs < 0 || !dist || !energy || !grad || !vids ||
      (kind == 0 && !tris) || (kind == 1 && !edges) || !(d_hat > 0.f))
    return fail(MFEM_EINVAL, "bad argument");
  if (max_pairs == 0) return MFEM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (kind == 0)
    hipLaunchKernelGGL(k_pair_barrier<0>, dim3(blocks(max_pairs)), dim3(kThreads), 0, st, x, pairs, n_pairs, max_pairs,
                       tris, edges, kappa, d_hat, dist, energy, grad, vids);
  else
    hipLaunchKernelGGL(k_pair_barrier<1>, dim3(blocks(max_pairs)), dim3(kThreads), 0, st, x, pairs, n_pairs, max_pairs,
                       tris, edges, kappa, d_hat, dist, energy, grad, vids);
  return launched();
}

}  // extern "C"
