// Level-synchronous reverse Cuthill-McKee on the device, bit-exact to the sequential
// 3DRCM order (InchingLiteInteger/Fuel/Coordinate/T1.py:171-323).
//
// The sequential algorithm appends, for each dequeued node in queue order, its unvisited
// neighbours (scanned in descending index order) and sorts those new children stably by
// degree. Queue order is level order, so level L+1 consists of the unvisited neighbours of
// level L, each one the child of the first level-L node (lowest queue position) adjacent to
// it, and inside one parent's group the children are ordered by (degree ascending, index
// descending). One level is built per round:
//   claim   : claim[j] = min position of a level-L neighbour           (atomicMin)
//   count   : cnt[p]   = number of nodes claimed by the node at position p
//   scan    : off      = exclusive prefix sum of cnt (single workgroup)
//   scatter : children written into their parent's slot range        (atomic slot)
//   sort    : each group sorted by (degree, -index) (one wave per parent, rank by key)
//   place   : pos[node] = its position
// The level size comes back to the host after the scan; a new component starts from the
// first unvisited node in the caller's seed order.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "common.h"

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) INCH_FAIL(INCH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

namespace {

constexpr int kG = 16;  // lanes per level node in the neighbour scans

__global__ void k_rcm_init(int64_t n, int32_t *__restrict__ pos, int32_t *__restrict__ claim) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    pos[i] = -1;
    claim[i] = INT_MAX;
  }
}

__global__ void k_rcm_zero(int32_t L, int32_t *__restrict__ a, int32_t *__restrict__ b) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < L) {
    a[i] = 0;
    b[i] = 0;
  }
}

__global__ void k_rcm_claim(int32_t s, int32_t e, const int32_t *__restrict__ ord, const int64_t *__restrict__ indptr,
                            const int32_t *__restrict__ indices, const int32_t *__restrict__ pos,
                            int32_t *__restrict__ claim) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t p = s + t / kG;
  const int lane = (int)(t % kG);
  if (p >= e) return;
  const int32_t i = ord[p];
  for (int64_t b = indptr[i] + lane; b < indptr[i + 1]; b += kG) {
    const int32_t j = indices[b];
    if (pos[j] < 0) atomicMin(&claim[j], (int32_t)p);
  }
}

__global__ void k_rcm_count(int32_t s, int32_t e, const int32_t *__restrict__ ord, const int64_t *__restrict__ indptr,
                            const int32_t *__restrict__ indices, const int32_t *__restrict__ pos,
                            const int32_t *__restrict__ claim, int32_t *__restrict__ cnt) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t p = s + t / kG;
  const int lane = (int)(t % kG);
  if (p >= e) return;
  const int32_t i = ord[p];
  int32_t c = 0;
  for (int64_t b = indptr[i] + lane; b < indptr[i + 1]; b += kG) {
    const int32_t j = indices[b];
    if (pos[j] < 0 && claim[j] == (int32_t)p) ++c;
  }
  if (c) atomicAdd(&cnt[p - s], c);
}

// off[0..L] = exclusive prefix sum of cnt[0..L); off[L] = total. One workgroup of 1024.
__global__ void __launch_bounds__(1024) k_rcm_scan(int32_t L, const int32_t *__restrict__ cnt,
                                                   int32_t *__restrict__ off) {
  __shared__ int32_t wsum[16];
  __shared__ int32_t carry_s;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) carry_s = 0;
  __syncthreads();
  for (int32_t base = 0; base < L; base += 1024) {
    const int32_t v = (base + t < L) ? cnt[base + t] : 0;
    int32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int32_t y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
      int32_t ws = lane < 16 ? wsum[lane] : 0;
#pragma unroll
      for (int d = 1; d < 16; d <<= 1) {
        const int32_t y = __shfl_up(ws, d, 64);
        if (lane >= d) ws += y;
      }
      if (lane < 16) wsum[lane] = ws;   // inclusive prefix of wave totals
    }
    __syncthreads();
    const int32_t carry = carry_s;
    const int32_t incl = x + (w > 0 ? wsum[w - 1] : 0);
    if (base + t < L) off[base + t] = carry + incl - v;
    __syncthreads();
    if (t == 1023) carry_s = carry + incl;
    __syncthreads();
  }
  if (t == 0) off[L] = carry_s;
}

__global__ void k_rcm_scatter(int32_t s, int32_t e, int32_t *__restrict__ ord, const int64_t *__restrict__ indptr,
                              const int32_t *__restrict__ indices, const int32_t *__restrict__ pos,
                              const int32_t *__restrict__ claim, const int32_t *__restrict__ off,
                              int32_t *__restrict__ fill) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t p = s + t / kG;
  const int lane = (int)(t % kG);
  if (p >= e) return;
  const int32_t i = ord[p];
  for (int64_t b = indptr[i] + lane; b < indptr[i + 1]; b += kG) {
    const int32_t j = indices[b];
    if (pos[j] < 0 && claim[j] == (int32_t)p) {
      const int32_t slot = atomicAdd(&fill[p - s], 1);
      ord[e + off[p - s] + slot] = j;
    }
  }
}

// Children of one parent: ascending degree, then descending index. One wave per parent:
// each child's key (degree << 32 | ~index) goes to LDS and its rank among the group's keys
// (all distinct) is its slot, which is the order a stable insertion sort by the same rule
// gives. Groups above kSortCap children (not produced by the neighbour lists, whose rows
// hold at most 2048 entries) are sorted by insertion in lane 0.
constexpr int kSortCap = 2048;

__global__ void __launch_bounds__(64) k_rcm_sort_groups(int32_t L, int32_t e, int32_t *__restrict__ ord,
                                                        const int64_t *__restrict__ indptr,
                                                        const int32_t *__restrict__ off) {
  __shared__ uint64_t key[kSortCap];
  const int32_t q = blockIdx.x;
  const int lane = threadIdx.x;
  if (q >= L) return;
  int32_t *g = ord + e + off[q];
  const int32_t m = off[q + 1] - off[q];
  if (m <= 1) return;
  if (m > kSortCap) {
    if (lane == 0) {
      for (int32_t a = 1; a < m; ++a) {
        const int32_t v = g[a];
        const int64_t dv = indptr[v + 1] - indptr[v];
        int32_t b = a;
        while (b > 0) {
          const int32_t u = g[b - 1];
          const int64_t du = indptr[u + 1] - indptr[u];
          if (du < dv || (du == dv && u > v)) break;
          g[b] = u;
          --b;
        }
        g[b] = v;
      }
    }
    return;
  }
  for (int32_t a = lane; a < m; a += 64) {
    const int32_t v = g[a];
    key[a] = ((uint64_t)(indptr[v + 1] - indptr[v]) << 32) | (uint64_t)(0xffffffffu - (uint32_t)v);
  }
  __syncthreads();
  for (int32_t a0 = 0; a0 < m; a0 += 64) {
    const int32_t a = a0 + lane;
    const uint64_t ka = a < m ? key[a] : 0;
    int32_t r = 0;
    for (int32_t b = 0; b < m; ++b) r += key[b] < ka;
    if (a < m) g[r] = (int32_t)(0xffffffffu - (uint32_t)ka);
  }
}

__global__ void k_rcm_place(int32_t e, int32_t L1, const int32_t *__restrict__ ord, int32_t *__restrict__ pos) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < L1) pos[ord[e + q]] = e + q;
}

// First k >= z with pos[seeds[k]] < 0 (atomicMin into *best, ULLONG_MAX when none). The
// block's minimum is formed first (wave shuffles, then the block's waves), so a block does
// at most one atomic and blocks with no unvisited seed none: with many components almost
// every entry past z is still unvisited, and one atomic per entry serialised on *best.
__global__ void __launch_bounds__(256) k_rcm_next_seed(int64_t n, int64_t z, const int64_t *__restrict__ seeds,
                                                       const int32_t *__restrict__ pos,
                                                       unsigned long long *__restrict__ best) {
  __shared__ unsigned long long wmin[4];
  const int64_t k = z + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  unsigned long long cand = (k < n && pos[seeds[k]] < 0) ? (unsigned long long)k : ULLONG_MAX;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(cand, off, 64);
    cand = o < cand ? o : cand;
  }
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = cand;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmin[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = wmin[w] < m ? wmin[w] : m;
    if (m != ULLONG_MAX) atomicMin(best, m);
  }
}

__global__ void k_rcm_start(int32_t N, const int64_t *__restrict__ seeds, const unsigned long long *__restrict__ best,
                            int32_t *__restrict__ ord, int32_t *__restrict__ pos) {
  const int32_t v = (int32_t)seeds[*best];
  ord[N] = v;
  pos[v] = N;
}

__global__ void k_rcm_finish(int64_t n, const int32_t *__restrict__ ord, int32_t reverse, int32_t *__restrict__ order,
                             int64_t *__restrict__ undo) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t v = reverse ? ord[n - 1 - p] : ord[p];
  order[p] = v;
  undo[v] = p;
}

__global__ void k_rcm_best_init(unsigned long long *best) { *best = ULLONG_MAX; }

// Input validation before any indexed access: seeds in [0, n), indptr non-decreasing from 0,
// columns in [0, n). Sets *bad to 1 on a violation.
__global__ void k_rcm_check(int64_t n, const int64_t *__restrict__ indptr, const int32_t *__restrict__ indices,
                            const int64_t *__restrict__ seeds, int32_t *__restrict__ bad) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) {
    const int64_t sd = seeds[t];
    if (sd < 0 || sd >= n) *bad = 1;
    if (indptr[t + 1] < indptr[t] || (t == 0 && indptr[0] != 0)) *bad = 1;
  }
}

__global__ void k_rcm_check_cols(int64_t nnz, int64_t n, const int32_t *__restrict__ indices,
                                 int32_t *__restrict__ bad) {
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nnz; b += (int64_t)gridDim.x * blockDim.x) {
    const int32_t j = indices[b];
    if (j < 0 || j >= n) *bad = 1;
  }
}

__global__ void k_rcm_flag_init(int32_t *bad) { *bad = 0; }

__global__ void k_rcm_seed_hist(int64_t n, const int64_t *__restrict__ seeds, int32_t *__restrict__ cnt) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) atomicAdd(&cnt[seeds[t]], 1);
}

__global__ void k_rcm_seed_perm(int64_t n, const int32_t *__restrict__ cnt, int32_t *__restrict__ bad) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n && cnt[t] != 1) *bad = 1;
}

unsigned blocks_for(int64_t threads, int bs) { return (unsigned)((threads + bs - 1) / bs); }

}  // namespace

extern "C" {

int64_t inch_rcm_device_work_size(int64_t n) { return 6 * (n + 2) * (int64_t)sizeof(int32_t) + 64; }

inch_status inch_rcm_device(int64_t n, const int64_t *indptr, const int32_t *indices, const int64_t *seeds,
                            int32_t reverse, int32_t *order, int64_t *undo, void *work, int64_t work_bytes,
                            int64_t *comp_ptr, int64_t comp_cap, int64_t *n_comp, void *stream) {
  INCH_REQUIRE(!comp_ptr || comp_cap >= 2, "inch_rcm_device: comp_cap must be >= 2");
  INCH_REQUIRE(n > 0 && n < INT_MAX - 2 && indptr && indices && seeds && order && undo && work,
               "inch_rcm_device: bad args");
  INCH_REQUIRE(work_bytes >= inch_rcm_device_work_size(n), "inch_rcm_device: work too small (%lld < %lld)",
               (long long)work_bytes, (long long)inch_rcm_device_work_size(n));
  hipStream_t st = (hipStream_t)stream;
  unsigned long long *best = (unsigned long long *)work;          // first 64 bytes: scalars
  int32_t *pos = (int32_t *)((char *)work + 64);
  int32_t *claim = pos + (n + 2);
  int32_t *cnt = claim + (n + 2);
  int32_t *off = cnt + (n + 2);
  int32_t *fill = off + (n + 2);
  int32_t *ord = fill + (n + 2);
  const int bs = 256;
  int32_t *bad = (int32_t *)((char *)work + 8);
  k_rcm_flag_init<<<1, 1, 0, st>>>(bad);
  k_rcm_check<<<blocks_for(n, bs), bs, 0, st>>>(n, indptr, indices, seeds, bad);
  int64_t nnz = 0;
  HIP_TRY(hipMemcpyAsync(&nnz, indptr + n, sizeof(nnz), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  INCH_REQUIRE(nnz >= 0, "inch_rcm_device: indptr[n] < 0");
  if (nnz > 0) k_rcm_check_cols<<<blocks_for(nnz < 4096 * 256 ? nnz : 4096 * 256, bs), bs, 0, st>>>(nnz, n, indices, bad);
  int32_t bad_h = 0;
  HIP_TRY(hipMemcpyAsync(&bad_h, bad, sizeof(bad_h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  INCH_REQUIRE(!bad_h, "inch_rcm_device: seeds, indptr or column indices out of range");
  k_rcm_zero<<<blocks_for(n, bs), bs, 0, st>>>((int32_t)n, cnt, fill);
  k_rcm_seed_hist<<<blocks_for(n, bs), bs, 0, st>>>(n, seeds, cnt);
  k_rcm_seed_perm<<<blocks_for(n, bs), bs, 0, st>>>(n, cnt, bad);
  HIP_TRY(hipMemcpyAsync(&bad_h, bad, sizeof(bad_h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  INCH_REQUIRE(!bad_h, "inch_rcm_device: seeds is not a permutation");
  k_rcm_init<<<blocks_for(n, bs), bs, 0, st>>>(n, pos, claim);
  HIP_TRY(hipGetLastError());
  int64_t N = 0, z = 0;
  std::vector<int64_t> starts;   // BFS position of each component's seed
  while (N < n) {
    if (z >= n) INCH_FAIL(INCH_ERR_INVALID, "inch_rcm_device: seeds is not a permutation");
    starts.push_back(N);
    k_rcm_best_init<<<1, 1, 0, st>>>(best);
    k_rcm_next_seed<<<blocks_for(n - z, bs), bs, 0, st>>>(n, z, seeds, pos, best);
    unsigned long long kbest = 0;
    HIP_TRY(hipMemcpyAsync(&kbest, best, sizeof(kbest), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (kbest == ULLONG_MAX) INCH_FAIL(INCH_ERR_INVALID, "inch_rcm_device: seeds is not a permutation");
    z = (int64_t)kbest + 1;
    k_rcm_start<<<1, 1, 0, st>>>((int32_t)N, seeds, best, ord, pos);
    int32_t s = (int32_t)N, e = (int32_t)N + 1;
    N += 1;
    while (true) {
      const int32_t L = e - s;
      k_rcm_zero<<<blocks_for(L, bs), bs, 0, st>>>(L, cnt, fill);
      k_rcm_claim<<<blocks_for((int64_t)L * kG, bs), bs, 0, st>>>(s, e, ord, indptr, indices, pos, claim);
      k_rcm_count<<<blocks_for((int64_t)L * kG, bs), bs, 0, st>>>(s, e, ord, indptr, indices, pos, claim, cnt);
      k_rcm_scan<<<1, 1024, 0, st>>>(L, cnt, off);
      int32_t L1 = 0;
      HIP_TRY(hipMemcpyAsync(&L1, off + L, sizeof(L1), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      if (L1 < 0 || N + L1 > n) INCH_FAIL(INCH_ERR_INVALID, "inch_rcm_device: inconsistent graph (level %d)", L1);
      if (L1 == 0) break;
      k_rcm_scatter<<<blocks_for((int64_t)L * kG, bs), bs, 0, st>>>(s, e, ord, indptr, indices, pos, claim, off,
                                                                    fill);
      k_rcm_sort_groups<<<(unsigned)L, 64, 0, st>>>(L, e, ord, indptr, off);
      k_rcm_place<<<blocks_for(L1, bs), bs, 0, st>>>(e, L1, ord, pos);
      HIP_TRY(hipGetLastError());
      s = e;
      e = e + L1;
      N += L1;
    }
  }
  k_rcm_finish<<<blocks_for(n, bs), bs, 0, st>>>(n, ord, reverse, order, undo);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  // components are contiguous in the output order: their start offsets, ascending, then n
  const int64_t C = (int64_t)starts.size();
  if (n_comp) *n_comp = C;
  if (comp_ptr) {
    if (C + 1 > comp_cap) INCH_FAIL(INCH_ERR_CAPACITY, "inch_rcm_device: %lld components exceed comp_cap", (long long)C);
    for (int64_t c = 0; c < C; ++c) {
      const int64_t s0 = starts[c], s1 = c + 1 < C ? starts[c + 1] : n;
      comp_ptr[reverse ? C - 1 - c : c] = reverse ? n - s1 : s0;
    }
    comp_ptr[C] = n;
  }
  return INCH_OK;
}

}  // extern "C"
