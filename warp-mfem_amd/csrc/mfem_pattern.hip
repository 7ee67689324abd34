// mfem_pattern.hip -- the system-matrix pattern and the deterministic
// assembly gather lists, built on the GPU after a topology change.
//
// The reference re-creates its BSR topology with warp.sparse
// (bsr_set_from_triplets, src/mfem/refinement/solver.py:685-715) after
// every refinement pass.  Here the SELL-64 pattern of H, the per-lane-slot
// lists of tet-block sources, the vertex -> (tet, corner) lists and the
// triangle-contact lists are rebuilt from the tets / triangles with sorts of
// 64-bit (major << 32 | minor) keys, so every list comes out in the same
// order as a sequential construction, whatever the scheduling:
//
//   adjacency : keys (i << 32 | j) of every vertex pair of every tet plus
//               (v << 32 | v) for every vertex, sorted, duplicates removed;
//               rowptr by binary search (lower bound of v << 32)
//   SELL-64   : width of a 64-row chunk = its longest row; chunk_ptr =
//               exclusive sum; row r's k-th entry goes to lane-slot
//               (chunk_ptr[r / 64] + k) * 64 + r % 64
//   contribs  : (lane-slot << 32 | source) per tet block, sorted
//   vinc      : (vertex << 32 | 4 t + corner), sorted
//   tcon/vtri : the same for triangle vertex pairs and corners
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "../../include/mfem_amd.h"

namespace {

constexpr int kThreads = 256;
thread_local std::string g_err;
int fail(int rc, const std::string& m) {
  g_err = m;
  return rc;
}
int launched() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MFEM_EHIP, hipGetErrorString(e));
  return MFEM_OK;
}
inline unsigned blocks(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// the 10 unique vertex pairs of a tet, as the assembly kernels number them
__constant__ int kPA[10] = {0, 1, 2, 3, 0, 0, 0, 1, 1, 2};
__constant__ int kPB[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};

__device__ __forceinline__ uint64_t pk(int hi, int lo) { return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo; }

// first index in sorted keys[0, n) with keys[i] >= x
__device__ __forceinline__ long lower_bound(const uint64_t* keys, long n, uint64_t x) {
  long lo = 0, hi = n;
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if (keys[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_pair_keys(const int32_t* tets, int n_tets, int n_verts, uint64_t* keys) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  const long n_pairs = 16L * n_tets;
  if (i < n_pairs) {
    const long t = i >> 4;
    const int a = (int)(i >> 2) & 3, b = (int)i & 3;
    keys[i] = pk(tets[4 * t + a], tets[4 * t + b]);
  } else if (i < n_pairs + n_verts) {
    const int v = (int)(i - n_pairs);
    keys[i] = pk(v, v);
  }
}

// rowptr[r] = lower bound of r << 32 in the unique keys (r <= n_verts)
__global__ void k_bounds(const uint64_t* keys, const int32_t* n_keys, int n_out, int32_t* ptr) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r > n_out) return;
  ptr[r] = (int32_t)lower_bound(keys, *n_keys, (uint64_t)(uint32_t)r << 32);
}
__global__ void k_bounds_n(const uint64_t* keys, long n_keys, int n_out, int32_t* ptr) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r > n_out) return;
  ptr[r] = (int32_t)lower_bound(keys, n_keys, (uint64_t)(uint32_t)r << 32);
}

__global__ void k_chunk_width(const int32_t* rowptr, int n_verts, int n_chunks, int32_t* width) {
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c > n_chunks) return;
  int w = 0;
  if (c < n_chunks)
    for (int r = 64 * c; r < 64 * c + 64 && r < n_verts; ++r) w = max(w, rowptr[r + 1] - rowptr[r]);
  width[c] = w;  // width[n_chunks] = 0: the exclusive sum's total lands there
}

__global__ void k_lanes(const uint64_t* keys, const int32_t* n_keys, const int32_t* rowptr, const int32_t* chunk_ptr,
                        int32_t* cols, int32_t* lane_of, int32_t* sell_col, int32_t* diag_lane) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= *n_keys) return;
  const int r = (int)(keys[i] >> 32), c = (int)(uint32_t)keys[i];
  const int lane = (chunk_ptr[r >> 6] + (int)(i - rowptr[r])) * 64 + (r & 63);
  cols[i] = c;
  lane_of[i] = lane;
  sell_col[lane] = c;
  if (c == r) diag_lane[r] = lane;
}

// lane-slot of block (a, b), or -1 when it is not in the pattern
__device__ __forceinline__ int find_lane(const int32_t* rowptr, const int32_t* cols, const int32_t* lane_of, int a,
                                         int b) {
  int lo = rowptr[a], hi = rowptr[a + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  return (lo < rowptr[a + 1] && cols[lo] == b) ? lane_of[lo] : -1;
}

// 16 (lane-slot << 32 | source) keys per tet: source = 2 (p T + t) for block
// (a, b) of pair p, + 1 for its transpose (b, a)
__global__ void k_contrib_keys(const int32_t* tets, int n_tets, const int32_t* rowptr, const int32_t* cols,
                               const int32_t* lane_of, uint64_t* keys) {
  const int t = blockIdx.x * kThreads + threadIdx.x;
  if (t >= n_tets) return;
  int o = 16 * t;
#pragma unroll
  for (int p = 0; p < 10; ++p) {
    const int ia = tets[4 * t + kPA[p]], ib = tets[4 * t + kPB[p]];
    const int src = 2 * (p * n_tets + t);
    keys[o++] = pk(find_lane(rowptr, cols, lane_of, ia, ib), src);
    if (kPA[p] != kPB[p]) keys[o++] = pk(find_lane(rowptr, cols, lane_of, ib, ia), src + 1);
  }
}

__global__ void k_corner_keys(const int32_t* idx, long n, uint64_t* keys) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) keys[i] = pk(idx[i], (int)i);
}

// 9 keys per triangle: (lane-slot << 32 | 9 f + 3 a + b)
__global__ void k_tri_keys(const int32_t* tris, int n_tris, const int32_t* rowptr, const int32_t* cols,
                           const int32_t* lane_of, uint64_t* keys, int32_t* status) {
  const int f = blockIdx.x * kThreads + threadIdx.x;
  if (f >= n_tris) return;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const int lane = find_lane(rowptr, cols, lane_of, tris[3 * f + a], tris[3 * f + b]);
      if (lane < 0) atomicOr(status, 1);
      keys[9 * f + 3 * a + b] = pk(lane < 0 ? 0 : lane, 9 * f + 3 * a + b);
    }
}

__global__ void k_low_words(const uint64_t* keys, long n, int32_t* out) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) out[i] = (int32_t)(uint32_t)keys[i];
}

__global__ void k_fill32(int32_t* p, long n, int32_t v) {
  const long i = (long)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) p[i] = v;
}

// workspace: two key arrays of the largest sort, the unique count, chunk
// widths, the lane of every pattern entry, and the library temporaries
struct PLayout {
  size_t off_a, off_b, off_cnt, off_width, off_lane, off_tmp, tmp_bytes, bytes;
};
PLayout p_layout(long max_verts, long max_tets, long max_tris) {
  PLayout L{};
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const long nk = std::max(16 * max_tets + max_verts, 9 * std::max(max_tris, 1L));
  const long n_chunks = (max_verts + 63) / 64;
  size_t o = 0;
  L.off_a = o; o += al((size_t)nk * 8);
  L.off_b = o; o += al((size_t)nk * 8);
  L.off_cnt = o; o += al(64);
  L.off_width = o; o += al((size_t)(n_chunks + 2) * 4);
  L.off_lane = o; o += al((size_t)(16 * max_tets + max_verts) * 4);
  size_t t_sort = 0, t_uniq = 0, t_scan = 0;
  (void)hipcub::DeviceRadixSort::SortKeys(nullptr, t_sort, (uint64_t*)nullptr, (uint64_t*)nullptr, (int)nk);
  (void)hipcub::DeviceSelect::Unique(nullptr, t_uniq, (uint64_t*)nullptr, (uint64_t*)nullptr, (int32_t*)nullptr, (int)nk);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan, (int32_t*)nullptr, (int32_t*)nullptr, (int)(n_chunks + 1));
  L.tmp_bytes = std::max(t_sort, std::max(t_uniq, t_scan));
  L.off_tmp = o; o += al(L.tmp_bytes);
  L.bytes = o;
  return L;
}

struct Ws {
  uint64_t *a, *b;
  int32_t* cnt;
  int32_t* width;
  int32_t* lane;
  void* tmp;
  size_t tmp_bytes;
};
Ws carve(void* ws, const PLayout& L) {
  char* p = static_cast<char*>(ws);
  return Ws{reinterpret_cast<uint64_t*>(p + L.off_a), reinterpret_cast<uint64_t*>(p + L.off_b),
            reinterpret_cast<int32_t*>(p + L.off_cnt), reinterpret_cast<int32_t*>(p + L.off_width),
            reinterpret_cast<int32_t*>(p + L.off_lane), p + L.off_tmp, L.tmp_bytes};
}

int sort_keys(Ws& W, uint64_t* in, uint64_t* out, long n, hipStream_t st) {
  size_t tb = W.tmp_bytes;
  if (n <= 0) return MFEM_OK;
  if (hipcub::DeviceRadixSort::SortKeys(W.tmp, tb, in, out, (int)n, 0, 64, st) != hipSuccess)
    return fail(MFEM_EHIP, "radix sort failed");
  return launched();
}

int check_args(int n_verts, int n_tets, int n_tris, int max_verts, int max_tets, int max_tris, const int32_t* tets,
               const int32_t* tris, void* ws, size_t ws_bytes, PLayout& L) {
  if (!tets || !ws) return fail(MFEM_EINVAL, "null argument");
  if (n_verts <= 0 || n_tets <= 0 || n_tris < 0 || n_verts > max_verts || n_tets > max_tets || n_tris > max_tris)
    return fail(MFEM_EINVAL, "bad counts");
  if (n_tris > 0 && !tris) return fail(MFEM_EINVAL, "null triangles");
  if (16L * max_tets + max_verts >= (1L << 31)) return fail(MFEM_ECAPACITY, "mesh too large for int32 lists");
  L = p_layout(max_verts, max_tets, max_tris);
  if (ws_bytes < L.bytes) return fail(MFEM_ECAPACITY, "pattern workspace too small");
  return MFEM_OK;
}

}  // namespace

extern "C" {

const char* mfem_pattern_last_error(void) { return g_err.c_str(); }

size_t mfem_pattern_workspace_bytes(int max_verts, int max_tets, int max_tris) {
  return p_layout(max_verts, max_tets, max_tris).bytes;
}

int mfem_pattern_count(const int32_t* tets, int n_tets, int n_verts, int max_verts, int max_tets, int max_tris,
                       void* workspace, size_t workspace_bytes, int32_t* rowptr, int32_t* chunk_ptr, int32_t* counts,
                       void* stream) {
  PLayout L;
  if (int rc = check_args(n_verts, n_tets, 0, max_verts, max_tets, max_tris, tets, nullptr, workspace,
                          workspace_bytes, L))
    return rc;
  if (!rowptr || !chunk_ptr || !counts) return fail(MFEM_EINVAL, "null output");
  hipStream_t st = (hipStream_t)stream;
  Ws W = carve(workspace, L);
  const long nk = 16L * n_tets + n_verts;
  const int n_chunks = (n_verts + 63) / 64;
  hipLaunchKernelGGL(k_pair_keys, dim3(blocks(nk)), dim3(kThreads), 0, st, tets, n_tets, n_verts, W.a);
  if (int rc = launched()) return rc;
  if (int rc = sort_keys(W, W.a, W.b, nk, st)) return rc;
  size_t tb = W.tmp_bytes;
  if (hipcub::DeviceSelect::Unique(W.tmp, tb, W.b, W.a, W.cnt, (int)nk, st) != hipSuccess)
    return fail(MFEM_EHIP, "unique failed");
  hipLaunchKernelGGL(k_bounds, dim3(blocks(n_verts + 1)), dim3(kThreads), 0, st, (const uint64_t*)W.a,
                     (const int32_t*)W.cnt, n_verts, rowptr);
  hipLaunchKernelGGL(k_chunk_width, dim3(blocks(n_chunks + 1)), dim3(kThreads), 0, st, (const int32_t*)rowptr,
                     n_verts, n_chunks, W.width);
  if (int rc = launched()) return rc;
  tb = W.tmp_bytes;
  if (hipcub::DeviceScan::ExclusiveSum(W.tmp, tb, W.width, chunk_ptr, n_chunks + 1, st) != hipSuccess)
    return fail(MFEM_EHIP, "scan failed");
  // counts = [nnz, n_slots]
  if (hipMemcpyAsync(counts, W.cnt, 4, hipMemcpyDeviceToDevice, st) != hipSuccess ||
      hipMemcpyAsync(counts + 1, chunk_ptr + n_chunks, 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return fail(MFEM_EHIP, "copy failed");
  return launched();
}

int mfem_pattern_fill(const int32_t* tets, int n_tets, const int32_t* tris, int n_tris, int n_verts, int max_verts,
                      int max_tets, int max_tris, void* workspace, size_t workspace_bytes, const int32_t* rowptr,
                      const int32_t* chunk_ptr, int n_slots, int32_t* cols, int32_t* sell_col, int32_t* diag_lane,
                      int32_t* contrib_ptr, int32_t* contrib_src, int32_t* vinc_ptr, int32_t* vinc_src,
                      int32_t* tcon_ptr, int32_t* tcon_src, int32_t* vtri_ptr, int32_t* vtri_src, int32_t* status,
                      void* stream) {
  PLayout L;
  if (int rc = check_args(n_verts, n_tets, n_tris, max_verts, max_tets, max_tris, tets, tris, workspace,
                          workspace_bytes, L))
    return rc;
  if (!rowptr || !chunk_ptr || !cols || !sell_col || !diag_lane || !contrib_ptr || !contrib_src || !vinc_ptr ||
      !vinc_src || !tcon_ptr || !vtri_ptr || !status || (n_tris > 0 && (!tcon_src || !vtri_src)))
    return fail(MFEM_EINVAL, "null output");
  if ((long)n_slots * 64 >= (1L << 31)) return fail(MFEM_ECAPACITY, "pattern too large for int32 lane ids");
  hipStream_t st = (hipStream_t)stream;
  Ws W = carve(workspace, L);
  const long n_lanes = (long)n_slots * 64;
  // (W.a holds the unique adjacency keys and W.cnt their count from
  // mfem_pattern_count)
  hipLaunchKernelGGL(k_fill32, dim3(blocks(n_lanes)), dim3(kThreads), 0, st, sell_col, n_lanes, 0);
  hipLaunchKernelGGL(k_fill32, dim3(1), dim3(kThreads), 0, st, status, 1L, 0);
  hipLaunchKernelGGL(k_lanes, dim3(blocks(16L * n_tets + n_verts)), dim3(kThreads), 0, st, (const uint64_t*)W.a,
                     (const int32_t*)W.cnt, rowptr, chunk_ptr, cols, W.lane, sell_col, diag_lane);
  if (int rc = launched()) return rc;
  // tet-block sources per lane-slot (W.a is free from here on)
  hipLaunchKernelGGL(k_contrib_keys, dim3(blocks(n_tets)), dim3(kThreads), 0, st, tets, n_tets, rowptr,
                     (const int32_t*)cols, (const int32_t*)W.lane, W.a);
  if (int rc = launched()) return rc;
  if (int rc = sort_keys(W, W.a, W.b, 16L * n_tets, st)) return rc;
  hipLaunchKernelGGL(k_low_words, dim3(blocks(16L * n_tets)), dim3(kThreads), 0, st, (const uint64_t*)W.b,
                     16L * n_tets, contrib_src);
  hipLaunchKernelGGL(k_bounds_n, dim3(blocks(n_lanes + 1)), dim3(kThreads), 0, st, (const uint64_t*)W.b,
                     16L * n_tets, (int)n_lanes, contrib_ptr);
  // vertex -> (tet, corner)
  hipLaunchKernelGGL(k_corner_keys, dim3(blocks(4L * n_tets)), dim3(kThreads), 0, st, tets, 4L * n_tets, W.a);
  if (int rc = launched()) return rc;
  if (int rc = sort_keys(W, W.a, W.b, 4L * n_tets, st)) return rc;
  hipLaunchKernelGGL(k_low_words, dim3(blocks(4L * n_tets)), dim3(kThreads), 0, st, (const uint64_t*)W.b,
                     4L * n_tets, vinc_src);
  hipLaunchKernelGGL(k_bounds_n, dim3(blocks(n_verts + 1)), dim3(kThreads), 0, st, (const uint64_t*)W.b,
                     4L * n_tets, n_verts, vinc_ptr);
  if (int rc = launched()) return rc;
  // triangle lists
  if (n_tris > 0) {
    hipLaunchKernelGGL(k_tri_keys, dim3(blocks(n_tris)), dim3(kThreads), 0, st, tris, n_tris, rowptr,
                       (const int32_t*)cols, (const int32_t*)W.lane, W.a, status);
    if (int rc = launched()) return rc;
    if (int rc = sort_keys(W, W.a, W.b, 9L * n_tris, st)) return rc;
    hipLaunchKernelGGL(k_low_words, dim3(blocks(9L * n_tris)), dim3(kThreads), 0, st, (const uint64_t*)W.b,
                       9L * n_tris, tcon_src);
    hipLaunchKernelGGL(k_bounds_n, dim3(blocks(n_lanes + 1)), dim3(kThreads), 0, st, (const uint64_t*)W.b,
                       9L * n_tris, (int)n_lanes, tcon_ptr);
    hipLaunchKernelGGL(k_corner_keys, dim3(blocks(3L * n_tris)), dim3(kThreads), 0, st, tris, 3L * n_tris, W.a);
    if (int rc = launched()) return rc;
    if (int rc = sort_keys(W, W.a, W.b, 3L * n_tris, st)) return rc;
    hipLaunchKernelGGL(k_low_words, dim3(blocks(3L * n_tris)), dim3(kThreads), 0, st, (const uint64_t*)W.b,
                       3L * n_tris, vtri_src);
    hipLaunchKernelGGL(k_bounds_n, dim3(blocks(n_verts + 1)), dim3(kThreads), 0, st, (const uint64_t*)W.b,
                       3L * n_tris, n_verts, vtri_ptr);
  } else {
    hipLaunchKernelGGL(k_fill32, dim3(blocks(n_lanes + 1)), dim3(kThreads), 0, st, tcon_ptr, n_lanes + 1, 0);
    hipLaunchKernelGGL(k_fill32, dim3(blocks(n_verts + 1)), dim3(kThreads), 0, st, vtri_ptr, (long)n_verts + 1, 0);
  }
  return launched();
}

}  // extern "C"
