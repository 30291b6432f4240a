// Host-side entry points: error state, synthetic globules, reverse Cuthill-McKee,
// row-batch partition. No GPU calls in this file.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace inch {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

// splitmix64 seeding + xoshiro256** generator (public-domain algorithms by
// S. Vigna); doubles in [0,1) from the top 53 bits.
struct Xoshiro {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  explicit Xoshiro(uint64_t seed) {
    uint64_t x = seed;
    for (auto &v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

}  // namespace inch

extern "C" {

const char *inch_status_string(inch_status s) {
  switch (s) {
    case INCH_OK: return "ok";
    case INCH_ERR_INVALID: return "invalid argument";
    case INCH_ERR_HIP: return "HIP runtime error";
    case INCH_ERR_CAPACITY: return "capacity exceeded";
    case INCH_ERR_NOT_CONVERGED: return "not converged";
    case INCH_ERR_NAN: return "NaN detected";
    default: return "unknown status";
  }
}

const char *inch_last_error(void) { return inch::g_last_error.c_str(); }

int32_t inch_abi_version(void) { return 1; }

inch_status inch_synth_globule(int64_t n, double radius, double dmin, uint64_t seed,
                               int64_t max_attempts, double *xyz) {
  INCH_REQUIRE(n >= 0 && radius > 0 && dmin >= 0 && xyz != nullptr, "inch_synth_globule: bad args");
  if (n == 0) return INCH_OK;
  inch::Xoshiro rng(seed);
  const double cell = dmin > 0 ? dmin : radius;
  const int64_t g = std::max<int64_t>(1, (int64_t)std::ceil(2.0 * radius / cell));
  if (g > 4096) INCH_FAIL(INCH_ERR_INVALID, "inch_synth_globule: grid too fine (%lld)", (long long)g);
  const int cap = 8;  // points with pairwise distance >= cell fit at most 8 per cube of side cell
  std::vector<int32_t> cnt((size_t)(g * g * g), 0);
  std::vector<int32_t> slot((size_t)(g * g * g) * cap, -1);
  const double d2min = dmin * dmin;
  int64_t placed = 0, attempts = 0;
  if (max_attempts <= 0) max_attempts = 1000 * n + 1000000;
  while (placed < n) {
    if (attempts++ >= max_attempts)
      INCH_FAIL(INCH_ERR_CAPACITY, "inch_synth_globule: saturated after %lld darts with %lld of %lld atoms",
                (long long)attempts, (long long)placed, (long long)n);
    double p[3];
    double r2 = 0;
    for (int a = 0; a < 3; ++a) {
      p[a] = (2.0 * rng.uniform() - 1.0) * radius;
      r2 += p[a] * p[a];
    }
    if (r2 > radius * radius) continue;
    int64_t c[3];
    for (int a = 0; a < 3; ++a) c[a] = std::min<int64_t>(g - 1, std::max<int64_t>(0, (int64_t)((p[a] + radius) / cell)));
    bool ok = true;
    for (int64_t dx = -1; dx <= 1 && ok; ++dx)
      for (int64_t dy = -1; dy <= 1 && ok; ++dy)
        for (int64_t dz = -1; dz <= 1 && ok; ++dz) {
          int64_t cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
          if (cx < 0 || cy < 0 || cz < 0 || cx >= g || cy >= g || cz >= g) continue;
          int64_t ci = (cx * g + cy) * g + cz;
          for (int s = 0; s < cnt[ci]; ++s) {
            const double *q = xyz + 3 * (int64_t)slot[ci * cap + s];
            double e0 = q[0] - p[0], e1 = q[1] - p[1], e2 = q[2] - p[2];
            if (e0 * e0 + e1 * e1 + e2 * e2 < d2min) { ok = false; break; }
          }
        }
    if (!ok) continue;
    int64_t ci = (c[0] * g + c[1]) * g + c[2];
    if (cnt[ci] >= cap) continue;
    slot[ci * cap + cnt[ci]++] = (int32_t)placed;
    for (int a = 0; a < 3; ++a) xyz[3 * placed + a] = p[a];
    ++placed;
  }
  return INCH_OK;
}

// Reverse Cuthill-McKee following the reference's 3DRCM (Fuel/Coordinate/T1.py:171-323):
// seeds in (stable) ascending-degree order, BFS visiting each node's neighbours in
// descending index order, the new children of a node stably sorted by degree, final
// reversal.
inch_status inch_rcm_host(int64_t n, const int64_t *indptr, const int32_t *indices,
                          const int64_t *seed_order, int32_t reverse, int32_t *order,
                          int64_t *undo) {
  INCH_REQUIRE(n >= 0 && indptr && order && undo && (n == 0 || indices), "inch_rcm_host: bad args");
  if (n == 0) return INCH_OK;
  std::vector<int32_t> deg((size_t)n);
  int32_t maxdeg = 0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t d = indptr[i + 1] - indptr[i];
    INCH_REQUIRE(d >= 0 && d < (1 << 30), "inch_rcm_host: bad indptr at %lld", (long long)i);
    deg[i] = (int32_t)d;
    maxdeg = std::max(maxdeg, deg[i]);
  }
  std::vector<int64_t> seeds((size_t)n);
  if (seed_order) {
    std::vector<char> seen((size_t)n, 0);
    for (int64_t i = 0; i < n; ++i) {
      int64_t s = seed_order[i];
      INCH_REQUIRE(s >= 0 && s < n && !seen[s], "inch_rcm_host: seed_order is not a permutation");
      seen[s] = 1;
      seeds[i] = s;
    }
  } else {
    // stable counting sort by degree
    std::vector<int64_t> start((size_t)maxdeg + 2, 0);
    for (int64_t i = 0; i < n; ++i) start[deg[i] + 1]++;
    for (int32_t d = 0; d <= maxdeg; ++d) start[d + 1] += start[d];
    for (int64_t i = 0; i < n; ++i) seeds[start[deg[i]]++] = i;
  }
  std::vector<char> visited((size_t)n, 0);
  std::vector<int32_t> ord((size_t)n);
  int64_t N = 0;
  for (int64_t zz = 0; zz < n && N < n; ++zz) {
    int64_t seed = seeds[zz];
    if (visited[seed]) continue;
    visited[seed] = 1;
    ord[N++] = (int32_t)seed;
    int64_t ls = N - 1, le = N;
    while (ls < le) {
      for (int64_t ii = ls; ii < le; ++ii) {
        int64_t i = ord[ii];
        int64_t n_old = N;
        for (int64_t p = indptr[i + 1] - 1; p >= indptr[i]; --p) {
          int32_t j = indices[p];
          if (j < 0 || j >= n) INCH_FAIL(INCH_ERR_INVALID, "inch_rcm_host: column %d out of range", j);
          if (!visited[j]) {
            visited[j] = 1;
            ord[N++] = j;
          }
        }
        // stable insertion sort of the new children by degree
        for (int64_t a = n_old + 1; a < N; ++a) {
          int32_t v = ord[a];
          int32_t dv = deg[v];
          int64_t b = a;
          while (b > n_old && dv < deg[ord[b - 1]]) {
            ord[b] = ord[b - 1];
            --b;
          }
          ord[b] = v;
        }
      }
      ls = le;
      le = N;
    }
  }
  for (int64_t p = 0; p < n; ++p) order[p] = reverse ? ord[n - 1 - p] : ord[p];
  for (int64_t p = 0; p < n; ++p) undo[order[p]] = p;
  return INCH_OK;
}

inch_status inch_batch_head(int64_t n, int64_t maxleaf, int64_t *head, int64_t cap, int64_t *nbatch) {
  INCH_REQUIRE(n >= 0 && maxleaf >= 1 && head && nbatch, "inch_batch_head: bad args");
  std::vector<int64_t> leaves, stack{n};
  while (!stack.empty()) {
    int64_t m = stack.back();
    stack.pop_back();
    if (m <= maxleaf) {
      leaves.push_back(m);
      continue;
    }
    int64_t k = m / 2;
    stack.push_back(m - (k + 1));
    stack.push_back(k + 1);
  }
  std::stable_sort(leaves.begin(), leaves.end(), [](int64_t a, int64_t b) { return a > b; });
  *nbatch = (int64_t)leaves.size();
  if ((int64_t)leaves.size() + 1 > cap) INCH_FAIL(INCH_ERR_CAPACITY, "inch_batch_head: need %lld heads", (long long)leaves.size() + 1);
  head[0] = 0;
  for (size_t b = 0; b < leaves.size(); ++b) head[b + 1] = head[b] + leaves[b];
  return INCH_OK;
}

}  // extern "C"
