// Level-synchronous reverse Cuthill-McKee on the device, bit-exact to the sequential
// 3DRCM order (InchingLiteInteger/Fuel/Coordinate/T1.py:171-323).
//
// The sequential algorithm appends, for each dequeued node in queue order, its unvisited
// neighbours (scanned in descending index order) and sorts those new children stably by
// degree; when the queue runs dry, the first unvisited node in the caller's seed order
// starts the next connected component. Queue order is level order, so level L+1 consists of
// the unvisited neighbours of level L, each one the child of the first level-L node (lowest
// queue position) adjacent to it, and inside one parent's group the children are ordered by
// (degree ascending, index descending). One level is built per round:
//   claim   : claim[j] = min position of a level-L neighbour           (atomicMin)
//   count   : cnt[p]   = number of nodes claimed by the node at position p
//   scan    : off      = exclusive prefix sum of cnt
//   scatter : children written into their parent's slot range        (atomic slot)
//   sort    : each group sorted by (degree, -index) (one wave per parent, rank by key)
//   place   : pos[node] = its position
// Components do not interact, so all of them advance in the same rounds: the components
// are found first (union-find), each one's seed is its first node in the seed order, and
// the components, ordered by their seeds' places in the seed order, are the order in which
// the sequential algorithm visits them. A round's frontier holds the components' levels one
// after another in that order, so inside each component positions compare as in the
// sequential run; every node's final position is its component's start plus its rank in
// the component (levels before its own, then its place in its own level). The number of
// rounds is the deepest component's level count, not the sum over components.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "common.h"
#include "scan.h"

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

// ---- components: union-find over the edges. A root is only ever hooked (CAS) onto a
// smaller root, so parent[v] <= v throughout and each component ends with its smallest node
// as root; finds shorten the paths they walk (the ECL-CC scheme, Jaiganesh and Burtscher,
// 2018). The finds' loads may see a stale parent: any value ever stored is an ancestor in
// the same component, and a hook only succeeds on a node that is a root at the CAS (which
// returns the current parent otherwise), so they are relaxed atomics of workgroup scope
// (cacheable); the final flattening runs after the hooking kernel has completed. ----
__device__ __forceinline__ int32_t cc_load(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void cc_store(int32_t *p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ int32_t cc_find(int32_t *parent, int32_t v) {
  int32_t par = cc_load(parent + v);
  if (par != v) {
    int32_t prev = v, next;
    while (par > (next = cc_load(parent + par))) {
      cc_store(parent + prev, next);
      prev = par;
      par = next;
    }
  }
  return par;
}

// Lanes per node of the two edge passes: a node's row is read by kCcLanes consecutive lanes
// (coalesced), not by one lane walking it alone.
constexpr int kCcLanes = 16;

// parent[v] = v's smallest neighbour when that is smaller than v (a valid hook: parent[v] <=
// v and in v's component), which leaves most trees shallow before any CAS
__global__ void k_cc_init(int64_t n, const int64_t *__restrict__ indptr, const int32_t *__restrict__ indices,
                          int32_t *__restrict__ parent) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t v = t / kCcLanes;
  const int lane = (int)(t % kCcLanes);
  if (v >= n) return;  // a node's lanes leave together
  int32_t m = (int32_t)v;
  for (int64_t b = indptr[v] + lane; b < indptr[v + 1]; b += kCcLanes) m = indices[b] < m ? indices[b] : m;
#pragma unroll
  for (int off = kCcLanes / 2; off > 0; off >>= 1) {
    const int32_t o = __shfl_xor(m, off, kCcLanes);
    m = o < m ? o : m;
  }
  if (lane == 0) parent[v] = m;
}

// Every edge once, from its larger end; a node's edges are split over its kCcLanes lanes,
// each with its own (possibly stale) root of v, which a failed CAS refreshes.
__global__ void k_cc_hook(int64_t n, const int64_t *__restrict__ indptr, const int32_t *__restrict__ indices,
                          int32_t *parent) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t v = t / kCcLanes;
  const int lane = (int)(t % kCcLanes);
  if (v >= n) return;
  int32_t vr = cc_find(parent, (int32_t)v);
  for (int64_t b = indptr[v] + lane; b < indptr[v + 1]; b += kCcLanes) {
    const int32_t u = indices[b];
    if (u >= v) continue;  // every edge once, from its larger end
    int32_t ur = cc_find(parent, u);
    while (vr != ur) {
      if (vr < ur) {
        const int32_t r = atomicCAS(parent + ur, ur, vr);
        if (r == ur) break;
        ur = r;  // ur was hooked meanwhile: continue from its new parent
      } else {
        const int32_t r = atomicCAS(parent + vr, vr, ur);
        if (r == vr) {
          vr = ur;
          break;
        }
        vr = r;
      }
    }
  }
}

// label[v] = root of v (in place over parent: writing a root only shortens paths)
__global__ void k_cc_flatten(int64_t n, int32_t *parent) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= n) return;
  int32_t r = (int32_t)v, p;
  while ((p = cc_load(parent + r)) != r) r = p;
  cc_store(parent + v, r);
}

// rank[seeds[k]] = k; minr[root] = smallest rank in the component (minr starts at INT_MAX)
__global__ void k_cc_rank(int64_t n, const int64_t *__restrict__ seeds, int32_t *__restrict__ rank,
                          int32_t *__restrict__ minr) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) {
    rank[seeds[k]] = (int32_t)k;
    minr[k] = INT_MAX;
  }
}

__global__ void k_cc_minrank(int64_t n, const int32_t *__restrict__ label, const int32_t *__restrict__ rank,
                             int32_t *__restrict__ minr) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v < n) atomicMin(&minr[label[v]], rank[v]);
}

// flag[k] = 1 when seeds[k] is the first node of its component in the seed order
__global__ void k_cc_flag(int64_t n, const int64_t *__restrict__ seeds, const int32_t *__restrict__ label,
                          const int32_t *__restrict__ minr, int32_t *__restrict__ flag) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < n) flag[k] = minr[label[seeds[k]]] == (int32_t)k ? 1 : 0;
}

// component c (in seed order) starts from S[c]; croot[root] = c; csize zeroed
__global__ void k_cc_seeds(int64_t n, const int64_t *__restrict__ seeds, const int32_t *__restrict__ label,
                           const int32_t *__restrict__ flag, const int32_t *__restrict__ cidx,
                           int32_t *__restrict__ S, int32_t *__restrict__ croot, int32_t *__restrict__ csize) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n) return;
  csize[k] = 0;
  if (flag[k]) {
    const int32_t v = (int32_t)seeds[k];
    S[cidx[k]] = v;
    croot[label[v]] = cidx[k];
  }
}

// cv[v] = component of v; csize[c] += 1 per node, one atomic per run of a wave's lanes that
// share a component (nodes of a component tend to be consecutive: few counters, heavy
// contention otherwise)
__global__ void __launch_bounds__(256) k_cc_vertex(int64_t n, const int32_t *__restrict__ label,
                                                   const int32_t *__restrict__ croot, int32_t *__restrict__ cv,
                                                   int32_t *__restrict__ csize) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int32_t c = v < n ? croot[label[v]] : -1;
  if (v < n) cv[v] = c;
  const int32_t cp = __shfl_up(c, 1, 64);
  const bool head = c >= 0 && (lane == 0 || cp != c);
  // run length: lanes up to the next head (or the wave's end)
  const uint64_t heads = __ballot(head);
  if (head) {
    const uint64_t later = lane < 63 ? heads >> (lane + 1) : 0;
    int run = later ? __ffsll((unsigned long long)later) : 64 - lane;
    if (v + run > n) run = (int)(n - v);
    atomicAdd(&csize[c], run);
  }
}

// round 0: the seeds, one per component, each first in its component
__global__ void k_rcm_seed_level(int32_t C, const int32_t *__restrict__ S, const int32_t *__restrict__ cstart,
                                 int32_t *__restrict__ ord, int32_t *__restrict__ pos, int32_t *__restrict__ fin,
                                 int32_t *__restrict__ done, int32_t *__restrict__ first,
                                 int32_t *__restrict__ lcnt) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  ord[c] = S[c];
  pos[S[c]] = c;
  fin[c] = cstart[c];
  done[c] = 1;
  first[c] = INT_MAX;
  lcnt[c] = 0;
}

// final positions of a new level [e, e + L1): its nodes of component c form one segment
// (the frontier is grouped by component, in component order, so their children are too),
// whose first node follows the done[c] nodes of c's earlier levels. Segment bounds come from
// comparing neighbouring positions: one writer each, no atomics.
__global__ void k_rcm_seg_first(int32_t e, int32_t L1, const int32_t *__restrict__ ord, const int32_t *__restrict__ cv,
                                int32_t *__restrict__ first, int32_t *__restrict__ lcnt) {
  const int32_t q = e + blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= e + L1) return;
  const int32_t c = cv[ord[q]];
  if (q == e || cv[ord[q - 1]] != c) first[c] = q;
  if (q == e + L1 - 1 || cv[ord[q + 1]] != c) lcnt[c] = q + 1;   // one past the segment's end
}

__global__ void k_rcm_seg_fin(int32_t e, int32_t L1, const int32_t *__restrict__ ord, const int32_t *__restrict__ cv,
                              const int32_t *__restrict__ cstart, const int32_t *__restrict__ done,
                              const int32_t *__restrict__ first, int32_t *__restrict__ fin) {
  const int32_t q = e + blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= e + L1) return;
  const int32_t c = cv[ord[q]];
  fin[q] = cstart[c] + done[c] + (q - first[c]);
}

__global__ void k_rcm_seg_close(int32_t e, int32_t L1, const int32_t *__restrict__ ord, const int32_t *__restrict__ cv,
                                int32_t *__restrict__ first, int32_t *__restrict__ lcnt, int32_t *__restrict__ done) {
  const int32_t q = e + blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= e + L1) return;
  const int32_t c = cv[ord[q]];
  if (first[c] == q) done[c] += lcnt[c] - q;
}

__global__ void k_rcm_finish(int64_t n, const int32_t *__restrict__ ord, const int32_t *__restrict__ fin,
                             int32_t reverse, int32_t *__restrict__ order, int64_t *__restrict__ undo) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int64_t p = reverse ? n - 1 - fin[q] : fin[q];
  const int32_t v = ord[q];
  order[p] = v;
  undo[v] = p;
}

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

constexpr int kRcmArrays = 20;  // int32 arrays of n + 2 entries in the work buffer

int64_t inch_rcm_device_work_size(int64_t n) {
  return (kRcmArrays * (n + 2) + 2 * (scan_tiles(n + 1) + 2)) * (int64_t)sizeof(int32_t) + 64;
}

inch_status inch_rcm_device(int64_t n, const int64_t *indptr, const int32_t *indices, const int64_t *seeds,
                            int32_t reverse, int32_t *order, int64_t *undo, void *work, int64_t work_bytes,
                            int64_t *comp_ptr, int64_t comp_cap, int64_t *n_comp, void *stream) {
  INCH_REQUIRE(!comp_ptr || comp_cap >= 2, "inch_rcm_device: comp_cap must be >= 2");
  INCH_REQUIRE(n > 0 && n < INT_MAX - 2 && indptr && indices && seeds && order && undo && work,
               "inch_rcm_device: bad args");
  INCH_REQUIRE(work_bytes >= inch_rcm_device_work_size(n), "inch_rcm_device: work too small (%lld < %lld)",
               (long long)work_bytes, (long long)inch_rcm_device_work_size(n));
  hipStream_t st = (hipStream_t)stream;
  int32_t *arr[kRcmArrays];
  for (int q = 0; q < kRcmArrays; ++q) arr[q] = (int32_t *)((char *)work + 64) + q * (n + 2);
  int32_t *pos = arr[0], *claim = arr[1], *cnt = arr[2], *off = arr[3], *fill = arr[4], *ord = arr[5];
  int32_t *label = arr[6], *rank = arr[7], *minr = arr[8], *flag = arr[9], *cidx = arr[10], *croot = arr[11];
  int32_t *cv = arr[12], *csize = arr[13], *cstart = arr[14], *fin = arr[15], *first = arr[16], *lcnt = arr[17];
  int32_t *done = arr[18], *S = arr[19];
  int32_t *tsum = arr[kRcmArrays - 1] + (n + 2), *toff = tsum + scan_tiles(n + 1) + 2;
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
  // components, their seeds (first node in the seed order) and sizes, in seed order
  k_cc_init<<<blocks_for(n * kCcLanes, bs), bs, 0, st>>>(n, indptr, indices, label);
  k_cc_hook<<<blocks_for(n * kCcLanes, bs), bs, 0, st>>>(n, indptr, indices, label);
  k_cc_flatten<<<blocks_for(n, bs), bs, 0, st>>>(n, label);
  k_cc_rank<<<blocks_for(n, bs), bs, 0, st>>>(n, seeds, rank, minr);
  k_cc_minrank<<<blocks_for(n, bs), bs, 0, st>>>(n, label, rank, minr);
  k_cc_flag<<<blocks_for(n, bs), bs, 0, st>>>(n, seeds, label, minr, flag);
  HIP_TRY(scan_exclusive_into<int32_t>(n, flag, cidx, tsum, toff, st));
  k_cc_seeds<<<blocks_for(n, bs), bs, 0, st>>>(n, seeds, label, flag, cidx, S, croot, csize);
  k_cc_vertex<<<blocks_for(n, bs), bs, 0, st>>>(n, label, croot, cv, csize);
  HIP_TRY(hipGetLastError());
  int32_t C = 0;
  HIP_TRY(hipMemcpyAsync(&C, cidx + n, sizeof(C), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (C < 1 || C > n) INCH_FAIL(INCH_ERR_INVALID, "inch_rcm_device: %d components", C);
  HIP_TRY(scan_exclusive_into<int32_t>(C, csize, cstart, tsum, toff, st));
  k_rcm_seed_level<<<blocks_for(C, bs), bs, 0, st>>>(C, S, cstart, ord, pos, fin, done, first, lcnt);
  HIP_TRY(hipGetLastError());
  // all components' levels advance together
  int32_t s = 0, e = C;
  int64_t N = C;
  while (true) {
    const int32_t L = e - s;
    k_rcm_zero<<<blocks_for(L, bs), bs, 0, st>>>(L, cnt, fill);
    k_rcm_claim<<<blocks_for((int64_t)L * kG, bs), bs, 0, st>>>(s, e, ord, indptr, indices, pos, claim);
    k_rcm_count<<<blocks_for((int64_t)L * kG, bs), bs, 0, st>>>(s, e, ord, indptr, indices, pos, claim, cnt);
    HIP_TRY(scan_exclusive_into<int32_t>(L, cnt, off, tsum, toff, st));
    int32_t L1 = 0;
    HIP_TRY(hipMemcpyAsync(&L1, off + L, sizeof(L1), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (L1 < 0 || N + L1 > n) INCH_FAIL(INCH_ERR_INVALID, "inch_rcm_device: inconsistent graph (level %d)", L1);
    if (L1 == 0) break;
    k_rcm_scatter<<<blocks_for((int64_t)L * kG, bs), bs, 0, st>>>(s, e, ord, indptr, indices, pos, claim, off,
                                                                  fill);
    k_rcm_sort_groups<<<(unsigned)L, 64, 0, st>>>(L, e, ord, indptr, off);
    k_rcm_place<<<blocks_for(L1, bs), bs, 0, st>>>(e, L1, ord, pos);
    k_rcm_seg_first<<<blocks_for(L1, bs), bs, 0, st>>>(e, L1, ord, cv, first, lcnt);
    k_rcm_seg_fin<<<blocks_for(L1, bs), bs, 0, st>>>(e, L1, ord, cv, cstart, done, first, fin);
    k_rcm_seg_close<<<blocks_for(L1, bs), bs, 0, st>>>(e, L1, ord, cv, first, lcnt, done);
    HIP_TRY(hipGetLastError());
    s = e;
    e = e + L1;
    N += L1;
  }
  if (N != n) INCH_FAIL(INCH_ERR_INVALID, "inch_rcm_device: %lld of %lld nodes ordered", (long long)N, (long long)n);
  k_rcm_finish<<<blocks_for(n, bs), bs, 0, st>>>(n, ord, fin, reverse, order, undo);
  HIP_TRY(hipGetLastError());
  std::vector<int32_t> cs;
  if (comp_ptr) {
    cs.resize((size_t)C + 1);
    HIP_TRY(hipMemcpyAsync(cs.data(), cstart, (C + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  }
  HIP_TRY(hipStreamSynchronize(st));
  // components are contiguous in the output order: their start offsets, ascending, then n
  if (n_comp) *n_comp = C;
  if (comp_ptr) {
    if (C + 1 > comp_cap) INCH_FAIL(INCH_ERR_CAPACITY, "inch_rcm_device: %d components exceed comp_cap", C);
    for (int64_t c = 0; c < C; ++c) {
      const int64_t s0 = cs[c], s1 = cs[c + 1];
      comp_ptr[reverse ? C - 1 - c : c] = reverse ? n - s1 : s0;
    }
    comp_ptr[C] = n;
  }
  return INCH_OK;
}

}  // extern "C"
