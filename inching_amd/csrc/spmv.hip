// Operator kernels for gfx950: full-storage 3x3 block-row SpMV / SpMM with fused
// epilogues (shift, Chebyshev three-term step, Hotelling term), the matrix-free
// FP64 ANM operator, and the FP64 dense kernels of the Lanczos / Rayleigh-Ritz
// steps (multi-dot, rank-J update, combine).
//
// Layout: row i of the block matrix is handled by a group of G lanes of one
// 64-lane wavefront (G in {4,8,16,32,64}); each lane streams every G-th 3x3 block
// of the row (36 B fp32 / 72 B fp64 contiguous per block, consecutive lanes read
// consecutive blocks) and gathers the 3 matching x entries; the group reduces its
// 3 partial sums with a fixed xor butterfly (deterministic) and lanes 0..2 of the
// group apply the epilogue to the row's 3 dofs.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.h"

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) INCH_FAIL(INCH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

namespace {

constexpr int kRedBlocks = 512;  // fixed grid of the two-phase reductions (determinism)

// Column indices: absolute int32 (CT = int32_t, rowbase = nullptr), or 16-bit offsets from
// the row's first column (CT = uint16_t, rowbase[row] = that column), 2 B per block instead
// of 4 B when every row of the RCM band spans fewer than 65,536 columns.
// XCD-aware row mapping: workgroups are dealt round-robin to the 8 XCDs, so workgroup w runs
// on XCD w % 8; remapping w to a contiguous range per XCD keeps each XCD's rows (and the x
// entries they gather) in one band, which its own L2 then holds.
__device__ __forceinline__ int64_t xcd_block(bool remap) {
  const int64_t bid = blockIdx.x;
  if (!remap) return bid;
  const int64_t nb = gridDim.x, q = nb >> 3, r = nb & 7, xcd = bid & 7, idx = bid >> 3;
  return xcd * q + (xcd < r ? xcd : r) + idx;   // a bijection of [0, nb)
}

template <typename MT, typename VT, int G, typename CT, bool XCD = false>
__global__ void __launch_bounds__(256) k_bsr_spmv(int64_t n, const int64_t *__restrict__ rowptr,
                                                   const CT *__restrict__ col, const int32_t *__restrict__ rowbase,
                                                   const MT *__restrict__ vals,
                                                   const VT *__restrict__ x, VT *__restrict__ y, MT s1, MT c0,
                                                   const VT *__restrict__ z, MT c2, VT *__restrict__ acc, MT mu,
                                                   const double *__restrict__ Q, const double *__restrict__ qcoef,
                                                   int mq) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t row = (xcd_block(XCD) * (int64_t)blockDim.x + threadIdx.x) / G;
  if (row >= n) return;  // whole groups leave together (a group never spans two rows)
  const int64_t b0 = rowptr[row], b1 = rowptr[row + 1];
  const int64_t cb = rowbase ? (int64_t)rowbase[row] : 0;
  MT a0 = 0, a1 = 0, a2 = 0;
  int64_t b = b0 + lane;
  for (; b + G < b1; b += 2 * G) {
    const int64_t ja = cb + (int64_t)col[b], jb = cb + (int64_t)col[b + G];
    const MT *va = vals + 9 * b;
    const MT *vb = vals + 9 * (b + G);
    MT xa0 = (MT)x[3 * ja], xa1 = (MT)x[3 * ja + 1], xa2 = (MT)x[3 * ja + 2];
    MT xb0 = (MT)x[3 * jb], xb1 = (MT)x[3 * jb + 1], xb2 = (MT)x[3 * jb + 2];
    a0 += va[0] * xa0 + va[1] * xa1 + va[2] * xa2;
    a1 += va[3] * xa0 + va[4] * xa1 + va[5] * xa2;
    a2 += va[6] * xa0 + va[7] * xa1 + va[8] * xa2;
    a0 += vb[0] * xb0 + vb[1] * xb1 + vb[2] * xb2;
    a1 += vb[3] * xb0 + vb[4] * xb1 + vb[5] * xb2;
    a2 += vb[6] * xb0 + vb[7] * xb1 + vb[8] * xb2;
  }
  if (b < b1) {
    const int64_t ja = cb + (int64_t)col[b];
    const MT *va = vals + 9 * b;
    MT xa0 = (MT)x[3 * ja], xa1 = (MT)x[3 * ja + 1], xa2 = (MT)x[3 * ja + 2];
    a0 += va[0] * xa0 + va[1] * xa1 + va[2] * xa2;
    a1 += va[3] * xa0 + va[4] * xa1 + va[5] * xa2;
    a2 += va[6] * xa0 + va[7] * xa1 + va[8] * xa2;
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off, G);
    a1 += __shfl_xor(a1, off, G);
    a2 += __shfl_xor(a2, off, G);
  }
  if (lane < 3) {
    const int64_t dof = 3 * row + lane;
    const MT w = lane == 0 ? a0 : (lane == 1 ? a1 : a2);
    MT out = w;
    if (c0 != (MT)0) out -= c0 * (MT)x[dof];
    out *= s1;
    if (z) out -= c2 * (MT)z[dof];
    if (qcoef) {
      double q = 0.0;
      const int64_t nd = 3 * n;
      for (int m = 0; m < mq; ++m) q += qcoef[m] * Q[m * nd + dof];
      out += (MT)q;
    }
    y[dof] = (VT)out;
    if (acc) acc[dof] = (VT)((MT)acc[dof] + mu * out);
  }
}

template <typename MT, typename VT, int G, int R>
__global__ void __launch_bounds__(256) k_bsr_spmm(int64_t n, const int64_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col, const MT *__restrict__ vals,
                                                   const VT *__restrict__ X, VT *__restrict__ Y, int nrhs) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const int r0 = blockIdx.y * R;
  if (row >= n) return;
  const int64_t b0 = rowptr[row], b1 = rowptr[row + 1];
  MT a[3][R];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < R; ++r) a[c][r] = 0;
  for (int64_t b = b0 + lane; b < b1; b += G) {
    const int64_t j = col[b];
    const MT *v = vals + 9 * b;
    MT h[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) h[t] = v[t];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r0 + r < nrhs) {
        MT x0 = (MT)X[(3 * j) * nrhs + r0 + r], x1 = (MT)X[(3 * j + 1) * nrhs + r0 + r],
           x2 = (MT)X[(3 * j + 2) * nrhs + r0 + r];
        a[0][r] += h[0] * x0 + h[1] * x1 + h[2] * x2;
        a[1][r] += h[3] * x0 + h[4] * x1 + h[5] * x2;
        a[2][r] += h[6] * x0 + h[7] * x1 + h[8] * x2;
      }
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < R; ++r) a[c][r] += __shfl_xor(a[c][r], off, G);
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r0 + r < nrhs) Y[(3 * row + c) * nrhs + r0 + r] = (VT)a[c][r];
  }
}

// 1/r2 in float64: the float64 hardware reciprocal (about 25 correct bits) refined by two
// Newton steps in FMA form, e = 1 - r2 inv, inv += inv e. The relative error of each step
// does not depend on the binade of r2: the seed's 2^-25 becomes about 2^-50 after the first
// step and stays below the final rounding after the second, so the result is within 1 ulp of
// the IEEE quotient 1.0 / r2 for every normal r2 (measured on gfx950 over 2^20 values of r2
// in [0.25, 100), six binades: equal in every case). Squared inter-atomic distances in A^2
// or nm^2 are never subnormal or near overflow, so no range reduction is needed.
__device__ __forceinline__ double recip_r2(double r2) {
  double inv = __builtin_amdgcn_rcp(r2);
  double e = __builtin_fma(-r2, inv, 1.0);
  inv = __builtin_fma(inv, e, inv);
  e = __builtin_fma(-r2, inv, 1.0);
  return __builtin_fma(inv, e, inv);
}

// One neighbour term of a cluster atom: a += d (d . (vi - xj)) / r2 with d = pj - pi, for a
// set list bit. The layout sets no bit for a pair at r2 = 0 (coincident atoms, masked as in
// the reference), so a set bit implies r2 > 0. Every rounding is fixed (explicit FMAs), so
// all forms of the cluster product add bitwise-equal terms. The whole term is one
// exec-masked branch, which a wave skips when none of its lanes has the bit.
__device__ __forceinline__ void cl_term(bool bit, const double (&pi)[3], const double (&vi)[3], double pj0,
                                        double pj1, double pj2, double xj0, double xj1, double xj2, double *a) {
  if (bit) {
    const double d0 = pj0 - pi[0], d1 = pj1 - pi[1], d2 = pj2 - pi[2];
    const double r2 = __builtin_fma(d2, d2, __builtin_fma(d1, d1, d0 * d0));
    const double w0 = vi[0] - xj0, w1 = vi[1] - xj1, w2 = vi[2] - xj2;
    const double s = __builtin_fma(d2, w2, __builtin_fma(d1, w1, d0 * w0)) * recip_r2(r2);
    a[0] = __builtin_fma(d0, s, a[0]);
    a[1] = __builtin_fma(d1, s, a[1]);
    a[2] = __builtin_fma(d2, s, a[2]);
  }
}

// y_i = plus_i x_i + sum_{j != i, r2 > 0} d (d . (x_i - x_j)) / r2,  d = p_j - p_i
// (the ANM product recomputed from float64 coordinates: H_ij = -d d^T / r2 and
// H_ii = -sum_j H_ij). Several right-hand sides, row-major [3n, nrhs]: blockIdx.y picks
// a chunk of R columns; the geometry of a neighbour is computed once per chunk.
template <int G, int R>
__global__ void __launch_bounds__(256) k_matfree(int64_t n, const double *__restrict__ xyz,
                                                  const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                  double plus_i, const double *__restrict__ X, double *__restrict__ Y,
                                                  int nrhs) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const int r0 = blockIdx.y * R;
  if (row >= n) return;
  const double pi0 = xyz[3 * row], pi1 = xyz[3 * row + 1], pi2 = xyz[3 * row + 2];
  double vi[3][R], a[3][R];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      vi[c][r] = (r0 + r < nrhs) ? X[(3 * row + c) * nrhs + r0 + r] : 0.0;
      a[c][r] = 0.0;
    }
  for (int64_t b = rowptr[row] + lane; b < rowptr[row + 1]; b += G) {
    const int64_t j = col[b];
    if (j == row) continue;
    const double d0 = xyz[3 * j] - pi0, d1 = xyz[3 * j + 1] - pi1, d2 = xyz[3 * j + 2] - pi2;
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (!(r2 > 0.0)) continue;
    const double inv = recip_r2(r2);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r0 + r < nrhs) {
        const double t = (d0 * (vi[0][r] - X[(3 * j) * nrhs + r0 + r]) + d1 * (vi[1][r] - X[(3 * j + 1) * nrhs + r0 + r]) +
                          d2 * (vi[2][r] - X[(3 * j + 2) * nrhs + r0 + r])) * inv;
        a[0][r] += d0 * t;
        a[1][r] += d1 * t;
        a[2][r] += d2 * t;
      }
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < R; ++r) a[c][r] += __shfl_xor(a[c][r], off, G);
  if (lane < 3) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r0 + r < nrhs) {
        const double w = lane == 0 ? a[0][r] : (lane == 1 ? a[1][r] : a[2][r]);
        const double v = lane == 0 ? vi[0][r] : (lane == 1 ? vi[1][r] : vi[2][r]);
        Y[(3 * row + lane) * nrhs + r0 + r] = w + plus_i * v;
      }
    }
  }
}

// Single right-hand side with the same fused epilogue as k_bsr_spmv:
//   out = s1 * (A x - c0 * x) - c2 * z (+ Q term); y = out; acc += mu * out.
// G lanes per row; each lane handles U neighbours per iteration with independent loads
// (memory-level parallelism for the dependent index -> coordinate/vector gathers). XCD = true
// gives each XCD a contiguous range of rows (xcd_block), so the band's gathers stay in its L2.
template <typename VT, int G, int U, typename CT, bool XCD = false>
__global__ void __launch_bounds__(256) k_matfree_epi(int64_t n, const double *__restrict__ xyz,
                                                      const int64_t *__restrict__ rowptr,
                                                      const CT *__restrict__ col,
                                                      const int32_t *__restrict__ rowbase, double plus_i,
                                                      const VT *__restrict__ x, VT *__restrict__ y, double s1,
                                                      double c0, const VT *__restrict__ z, double c2,
                                                      VT *__restrict__ acc, double mu, const double *__restrict__ Q,
                                                      const double *__restrict__ qcoef, int mq) {
  const int lane = threadIdx.x & (G - 1);
  const int64_t row = (xcd_block(XCD) * (int64_t)blockDim.x + threadIdx.x) / G;
  if (row >= n) return;
  const double pi0 = xyz[3 * row], pi1 = xyz[3 * row + 1], pi2 = xyz[3 * row + 2];
  const double vi0 = (double)x[3 * row], vi1 = (double)x[3 * row + 1], vi2 = (double)x[3 * row + 2];
  double a0 = 0, a1 = 0, a2 = 0;
  const int64_t b0 = rowptr[row], b1 = rowptr[row + 1];
  const int64_t cb = rowbase ? (int64_t)rowbase[row] : 0;   // 16-bit column offsets: row base
  for (int64_t b = b0 + lane; b < b1; b += G * U) {
    int64_t jj[U];
#pragma unroll
    for (int q = 0; q < U; ++q) jj[q] = (b + q * G < b1) ? cb + (int64_t)col[b + q * G] : row;
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int64_t j = jj[q];
      const double d0 = xyz[3 * j] - pi0, d1 = xyz[3 * j + 1] - pi1, d2 = xyz[3 * j + 2] - pi2;
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const double w0 = vi0 - (double)x[3 * j], w1 = vi1 - (double)x[3 * j + 1], w2 = vi2 - (double)x[3 * j + 2];
      // j == row gives d = 0 (no contribution); r2 == 0 is masked as in the reference
      const double t = r2 > 0.0 ? (d0 * w0 + d1 * w1 + d2 * w2) * recip_r2(r2) : 0.0;
      a0 += d0 * t;
      a1 += d1 * t;
      a2 += d2 * t;
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off, G);
    a1 += __shfl_xor(a1, off, G);
    a2 += __shfl_xor(a2, off, G);
  }
  if (lane < 3) {
    const int64_t dof = 3 * row + lane;
    const double vi = lane == 0 ? vi0 : (lane == 1 ? vi1 : vi2);
    double out = (lane == 0 ? a0 : (lane == 1 ? a1 : a2)) + plus_i * vi;
    if (c0 != 0.0) out -= c0 * vi;
    out *= s1;
    if (z) out -= c2 * (double)z[dof];
    if (qcoef) {
      double q = 0.0;
      const int64_t nd = 3 * n;
      for (int m = 0; m < mq; ++m) q += qcoef[m] * Q[m * nd + dof];
      out += q;
    }
    y[dof] = (VT)out;
    if (acc) acc[dof] = (VT)((double)acc[dof] + mu * out);
  }
}

// Cluster form of the matrix-free operator (layout from inch_anm_clusters): cluster c holds
// 4 atoms that are close in space and lists the union of their neighbours once as
// ((j - base[c]) << 8) | mask. One wave per cluster; the 64 lanes take 64 different list
// entries, so one load instruction gathers 64 neighbours, each used for the 4 atoms (the
// per-row kernel gathers every neighbour once per atom). The cluster's own coordinates and
// vector entries are wave-uniform (scalar registers). Each lane keeps 4 x 3 partial sums,
// reduced across the wave by halving exchanges in a fixed order. Per pair the arithmetic is
// that of k_matfree_epi; only the order of the sum over neighbours differs.
constexpr int kClAtoms = 4;

// One halving step across lane halves of width W (32 or 16) for a pair of doubles (a: lower
// half's value, b: upper half's value): v_permlane{W}_swap exchanges the upper W lanes of a
// with the lower W lanes of b, so a' + b' is, in every lane, the lane's half summed with
// its partner's copy of it (lower lanes: a of both, upper lanes: b of both).
template <int W>
__device__ __forceinline__ double swap_sum(double a, double b) {
  const uint32_t alo = (uint32_t)__double2loint(a), ahi = (uint32_t)__double2hiint(a);
  const uint32_t blo = (uint32_t)__double2loint(b), bhi = (uint32_t)__double2hiint(b);
  if constexpr (W == 32) {
    const auto l = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
  } else {
    const auto l = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
  }
}

// Sum over the wave's 64 lanes of a cluster's 12 per-lane partial sums, in a fixed order:
// two halving exchanges (lanes l and l ^ 32, then l and l ^ 16) leave 3 values per lane
// (value m + 3 (l / 16) in lane l, summed over l % 16 + 16 k), and 12 lanes add one value
// each over a 16-lane row through LDS. Lane m < 12 returns value m (atom m / 3, component
// m % 3); red: 3 x 64 doubles of LDS.
__device__ __forceinline__ double cl_reduce(const double (&v)[12], int lane, double (*red)[64]) {
  double u[6], r[3];
#pragma unroll
  for (int m = 0; m < 6; ++m) u[m] = swap_sum<32>(v[m], v[m + 6]);
#pragma unroll
  for (int m = 0; m < 3; ++m) r[m] = swap_sum<16>(u[m], u[m + 3]);
#pragma unroll
  for (int m = 0; m < 3; ++m) red[m][lane] = r[m];
  __syncthreads();
  double sum = 0.0;
  if (lane < 12) {
    const int row = lane / 3, m = lane % 3;
#pragma unroll
    for (int q = 0; q < 16; ++q) sum += red[m][16 * row + q];
  }
  __syncthreads();
  return sum;
}

// Raw buffer loads with 32-bit byte offsets from a descriptor built per array (word 3 of
// the gfx9-family descriptor is 0x00020000): no 64-bit address arithmetic per gather, and
// a load at or past num_records bytes returns 0 (hardware range check).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

// three consecutive VT at byte offset off, widened to double
__device__ __forceinline__ void load3(__amdgpu_buffer_rsrc_t r, uint32_t off, const double *, double (&o)[3]) {
  const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
  const auto b = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off + 16), 0, 0);
  o[0] = __hiloint2double((int)a[1], (int)a[0]);
  o[1] = __hiloint2double((int)a[3], (int)a[2]);
  o[2] = __hiloint2double((int)b[1], (int)b[0]);
}
__device__ __forceinline__ void load3(__amdgpu_buffer_rsrc_t r, uint32_t off, const float *, double (&o)[3]) {
  const auto a = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
  const auto b = __builtin_amdgcn_raw_buffer_load_b32(r, (int)(off + 8), 0, 0);
  o[0] = (double)__uint_as_float(a[0]);
  o[1] = (double)__uint_as_float(a[1]);
  o[2] = (double)__uint_as_float(b);
}

// The float64 single-entry form fits 80 VGPRs (6 waves per SIMD) without spilling.
template <typename VT, int U, bool XCD>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(U == 1 && sizeof(VT) == 8 ? 6 : 1)))
k_matfree_cl(int64_t n, const double *__restrict__ xyz, int64_t ncl, const int32_t *__restrict__ cluster_atoms,
             const int64_t *__restrict__ cl_ptr, const int32_t *__restrict__ base,
             const uint32_t *__restrict__ entries, double plus_i, const VT *__restrict__ x, VT *__restrict__ y,
             double s1, double c0, const VT *__restrict__ z, double c2, VT *__restrict__ acc, double mu,
             const double *__restrict__ Q, const double *__restrict__ qcoef, int mq) {
  const int64_t c = xcd_block(XCD);
  if (c >= ncl) return;
  const int lane = threadIdx.x;
  double pi[kClAtoms][3], vi[kClAtoms][3];
  __shared__ double red[3][64];  // cross-lane reduction
  double v[3 * kClAtoms];        // 4 atoms x 3 partial sums
#pragma unroll
  for (int t = 0; t < kClAtoms; ++t) {
    const int32_t at = cluster_atoms[kClAtoms * c + t];
    const int64_t i = at >= 0 ? at : 0;  // a padding slot has no list bits
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      pi[t][k] = xyz[3 * i + k];
      vi[t][k] = (double)x[3 * i + k];
    }
  }
#pragma unroll
  for (int q = 0; q < 3 * kClAtoms; ++q) v[q] = 0.0;
  // the epilogue's own loads depend only on the cluster's atoms: issue them before the list
  // (lane m < 12 writes value m: atom m / 3, component m % 3)
  const int32_t at_w = lane < 3 * kClAtoms ? cluster_atoms[kClAtoms * c + lane / 3] : -1;
  const int64_t dof = 3 * (int64_t)(at_w >= 0 ? at_w : 0) + lane % 3;
  double xv = 0.0, zv = 0.0, av = 0.0;
  if (at_w >= 0) {
    xv = (double)x[dof];
    if (z) zv = (double)z[dof];
    if (acc) av = (double)acc[dof];
  }
  const int64_t estart = cl_ptr[c];
  const uint32_t len = (uint32_t)(cl_ptr[c + 1] - estart);  // < 4096 (inch_anm_clusters)
  const uint32_t cb = (uint32_t)base[c];
  // 32-bit offsets: the cluster's own list (entries past it read as 0 = no term), and the
  // coordinate / vector rows (24 n and 3 n sizeof(VT) bytes < 2^32, inch_anm_clusters)
  const __amdgpu_buffer_rsrc_t rl = buf_rsrc(entries + estart, 4u * len);
  const __amdgpu_buffer_rsrc_t rp = buf_rsrc(xyz, (uint32_t)(24 * n));
  const __amdgpu_buffer_rsrc_t rx = buf_rsrc(x, (uint32_t)(3 * sizeof(VT) * n));
  // the next iteration's list entries are loaded while this iteration's gathers are in flight
  uint32_t wn[U];
#pragma unroll
  for (int q = 0; q < U; ++q) wn[q] = __builtin_amdgcn_raw_buffer_load_b32(rl, (int)(4u * (q * 64 + lane)), 0, 0);
  for (uint32_t e0 = 0; e0 < len; e0 += 64 * U) {
    uint32_t w[U];
    double pj[U][3], xj[U][3];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      w[q] = wn[q];
      const uint32_t j = cb + (w[q] >> 8);
      load3(rp, 24u * j, xyz, pj[q]);
      load3(rx, (uint32_t)(3 * sizeof(VT)) * j, x, xj[q]);
    }
#pragma unroll
    for (int q = 0; q < U; ++q)
      wn[q] = __builtin_amdgcn_raw_buffer_load_b32(rl, (int)(4u * (e0 + 64 * U + q * 64 + lane)), 0, 0);
    // all of this iteration's gathers and the next entries are in flight before any use
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const double pj0 = pj[q][0], pj1 = pj[q][1], pj2 = pj[q][2];
      const double xj0 = xj[q][0], xj1 = xj[q][1], xj2 = xj[q][2];
#pragma unroll
      for (int t = 0; t < kClAtoms; ++t)
        cl_term((w[q] >> t) & 1u, pi[t], vi[t], pj0, pj1, pj2, xj0, xj1, xj2, v + 3 * t);
    }
  }
  const double sum = cl_reduce(v, lane, red);
  if (at_w < 0) return;  // lanes >= 12 and padding slots
  double out = sum + plus_i * xv;
  if (c0 != 0.0) out -= c0 * xv;
  out *= s1;
  if (z) out -= c2 * zv;
  if (qcoef) {
    double qs = 0.0;
    const int64_t nd = 3 * n;
    for (int m = 0; m < mq; ++m) qs += qcoef[m] * Q[m * nd + dof];
    out += qs;
  }
  y[dof] = (VT)out;
  if (acc) acc[dof] = (VT)(av + mu * out);
}

// Persistent form of the cluster product: one wave takes cpw (<= 16) consecutive clusters.
// Their list pointers, bases and atom slots are loaded once (lane l holds the pointer and
// base of the wave's cluster l and its atom slot l), and the first list entries of the next
// cluster are loaded while the last list chunk of the current one is computed. A cluster's
// coordinates and vector rows (scalar loads) and its first neighbour gathers are then in
// flight together, so a wave waits on memory once per list chunk; the per-cluster form
// waits three times before its first chunk (list pointer, list entries, gathers). The
// clusters' sums wait in LDS and the epilogue runs once for all of them after the loop,
// 64 values per pass. Per lane the terms are added in the order of k_matfree_cl and the
// reduction is the same, so the two forms give bitwise-equal results.
constexpr int kClpMax = 16;  // clusters per wave

template <typename VT, bool XCD>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(VT) == 8 ? 6 : 1)))
k_matfree_clp(int64_t n, const double *__restrict__ xyz, int64_t ncl, const int32_t *__restrict__ cluster_atoms,
              const int64_t *__restrict__ cl_ptr, const int32_t *__restrict__ base,
              const uint32_t *__restrict__ entries, double plus_i, const VT *__restrict__ x, VT *__restrict__ y,
              double s1, double c0, const VT *__restrict__ z, double c2, VT *__restrict__ acc, double mu,
              const double *__restrict__ Q, const double *__restrict__ qcoef, int mq, int cpw) {
  const int64_t cw = xcd_block(XCD) * (int64_t)cpw;
  if (cw >= ncl) return;
  const int nc = (int)(ncl - cw < cpw ? ncl - cw : cpw);
  const int lane = threadIdx.x;
  __shared__ double red[3][64];
  __shared__ double sums[kClpMax * 3 * kClAtoms];  // value m of the wave's cluster ci at 12 ci + m
  // the wave's header: lane l <= nc holds cl_ptr[cw + l], lane l < nc base[cw + l],
  // lane l < 4 nc the atom of slot l (clusters cw .. cw + nc - 1)
  const int64_t hp = lane <= nc ? cl_ptr[cw + lane] : 0;
  const int32_t hb = lane < nc ? base[cw + lane] : 0;
  const int32_t ha = lane < kClAtoms * nc ? cluster_atoms[kClAtoms * cw + lane] : -1;
  const int hp_lo = (int)(uint32_t)hp, hp_hi = (int)(uint32_t)((uint64_t)hp >> 32);
  auto ptr_at = [&](int l) -> int64_t {
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane(hp_hi, l) << 32) |
                     (uint64_t)(uint32_t)__builtin_amdgcn_readlane(hp_lo, l));
  };
  const __amdgpu_buffer_rsrc_t rp = buf_rsrc(xyz, (uint32_t)(24 * n));
  const __amdgpu_buffer_rsrc_t rx = buf_rsrc(x, (uint32_t)(3 * sizeof(VT) * n));
  // the wave's lists are one range of at most 16 x 4096 entries: 32-bit offsets from its
  // start; an offset past the range (kOut) reads 0, an entry without a term
  const int64_t wstart = ptr_at(0);
  const __amdgpu_buffer_rsrc_t rw = buf_rsrc(entries + wstart, (uint32_t)(4 * (ptr_at(nc) - wstart)));
  constexpr uint32_t kOut = 0xfffffff0u;
  uint32_t off = 0, len = (uint32_t)(ptr_at(1) - wstart);  // len < 4096 (inch_anm_clusters)
  uint32_t wn = __builtin_amdgcn_raw_buffer_load_b32(rw, (int)((uint32_t)lane < len ? 4u * lane : kOut), 0, 0);
  for (int ci = 0; ci < nc; ++ci) {
    const uint32_t cb = (uint32_t)__builtin_amdgcn_readlane(hb, ci);
    // the next cluster's list (empty after the wave's last cluster)
    const uint32_t off_n = off + len;
    const uint32_t len_n = ci + 1 < nc ? (uint32_t)(ptr_at(ci + 2) - wstart) - off_n : 0u;
    double pi[kClAtoms][3], vi[kClAtoms][3];
#pragma unroll
    for (int t = 0; t < kClAtoms; ++t) {
      const int32_t at = __builtin_amdgcn_readlane(ha, kClAtoms * ci + t);
      const int64_t i = at >= 0 ? at : 0;  // a padding slot has no list bits
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        pi[t][k] = xyz[3 * i + k];
        vi[t][k] = (double)x[3 * i + k];
      }
    }
    double v[3 * kClAtoms];
#pragma unroll
    for (int q = 0; q < 3 * kClAtoms; ++q) v[q] = 0.0;
    for (uint32_t e0 = 0; e0 < len; e0 += 64) {
      const uint32_t w = wn;
      const uint32_t j = cb + (w >> 8);
      double pj[3], xj[3];
      load3(rp, 24u * j, xyz, pj);
      load3(rx, (uint32_t)(3 * sizeof(VT)) * j, x, xj);
      if (e0 + 64 < len) {
        const uint32_t q = e0 + 64 + lane;
        wn = __builtin_amdgcn_raw_buffer_load_b32(rw, (int)(q < len ? 4u * (off + q) : kOut), 0, 0);
      } else {
        wn = __builtin_amdgcn_raw_buffer_load_b32(rw, (int)((uint32_t)lane < len_n ? 4u * (off_n + lane) : kOut),
                                                  0, 0);
      }
      // the gathers and the next entries are in flight before any use
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < kClAtoms; ++t)
        cl_term((w >> t) & 1u, pi[t], vi[t], pj[0], pj[1], pj[2], xj[0], xj[1], xj[2], v + 3 * t);
    }
    if (len == 0)
      wn = __builtin_amdgcn_raw_buffer_load_b32(rw, (int)((uint32_t)lane < len_n ? 4u * (off_n + lane) : kOut), 0, 0);
    const double sum = cl_reduce(v, lane, red);
    if (lane < 3 * kClAtoms) sums[3 * kClAtoms * ci + lane] = sum;
    off = off_n;
    len = len_n;
  }
  __syncthreads();
  // epilogue: value q = 12 ci + m of the wave is atom slot 4 ci + m / 3, component m % 3
  const int64_t nd = 3 * n;
  for (int q0 = 0; q0 < 3 * kClAtoms * nc; q0 += 64) {
    const int q = q0 + lane;
    const int ci = q / (3 * kClAtoms), m = q % (3 * kClAtoms);
    const int32_t at = __builtin_amdgcn_ds_bpermute(4 * (kClAtoms * ci + m / 3), ha);
    if (q >= 3 * kClAtoms * nc || at < 0) continue;
    const int64_t dof = 3 * (int64_t)at + m % 3;
    const double xv = (double)x[dof];
    double out = sums[q] + plus_i * xv;
    if (c0 != 0.0) out -= c0 * xv;
    out *= s1;
    if (z) out -= c2 * (double)z[dof];
    if (qcoef) {
      double qs = 0.0;
      for (int mm = 0; mm < mq; ++mm) qs += qcoef[mm] * Q[mm * nd + dof];
      out += qs;
    }
    y[dof] = (VT)out;
    if (acc) acc[dof] = (VT)((double)acc[dof] + mu * out);
  }
}

// Record form of the cluster product, for the Chebyshev recurrence. The operator keeps its
// recurrence vectors in records rec[6 q .. 6 q + 5] = (coordinates, vector entries) of slot
// q = 4 c + t (atom t of cluster c; a padding slot holds zeros) and the lists as record byte
// offsets, 48 q_j | mask (the offset's low 4 bits are zero), ascending q_j
// (inch_anm_records). A neighbour is then one 48-byte record read by three 16-byte loads
// at the entry with its mask cleared, and a neighbouring cluster's four atoms are
// adjacent records: on the 2.4M-atom assembly this takes the product from 0.79 to 0.57 ms
// (gathers from the coordinate and vector arrays in the matrix's order touch about twice
// the cache lines). The epilogue is that of k_matfree_cl: rout's vector part = s1 (A x +
// (plus_i - c0) x) - c2 z with x and z the vector parts of rin and rz, acc += mu rout;
// the result also goes to yout (slot order, 12 per cluster) when given. rout's coordinate
// part is not written (the caller's buffers carry the coordinates).
template <bool XCD>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6)))
k_matfree_rec(int64_t ncl, const double *__restrict__ rin, const double *__restrict__ rz, double *__restrict__ rout,
              double *__restrict__ yout, const int64_t *__restrict__ cl_ptr, const uint32_t *__restrict__ entries,
              double plus_i, double s1, double c0, double c2, double *__restrict__ acc, double mu) {
  const int64_t c = xcd_block(XCD);
  if (c >= ncl) return;
  const int lane = threadIdx.x;
  __shared__ double red[3][64];
  double pi[kClAtoms][3], vi[kClAtoms][3];
#pragma unroll
  for (int t = 0; t < kClAtoms; ++t)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      pi[t][k] = rin[24 * c + 6 * t + k];
      vi[t][k] = rin[24 * c + 6 * t + 3 + k];
    }
  // lane m < 12 writes value m: slot 4 c + m / 3, component m % 3
  const int64_t o = 24 * c + 6 * (lane / 3) + 3 + lane % 3;
  double xv = 0.0, zv = 0.0, av = 0.0;
  if (lane < 3 * kClAtoms) {
    xv = rin[o];
    if (rz) zv = rz[o];
    if (acc) av = acc[12 * c + lane];
  }
  double v[3 * kClAtoms];
#pragma unroll
  for (int q = 0; q < 3 * kClAtoms; ++q) v[q] = 0.0;
  const int64_t estart = cl_ptr[c];
  const uint32_t len = (uint32_t)(cl_ptr[c + 1] - estart);  // < 4096 (inch_anm_clusters)
  const __amdgpu_buffer_rsrc_t rl = buf_rsrc(entries + estart, 4u * len);
  const __amdgpu_buffer_rsrc_t rr = buf_rsrc(rin, (uint32_t)(192 * ncl));  // 48 bytes per slot < 2^32
  uint32_t wn = __builtin_amdgcn_raw_buffer_load_b32(rl, (int)(4u * lane), 0, 0);
  for (uint32_t e0 = 0; e0 < len; e0 += 64) {
    const uint32_t w = wn;
    const uint32_t off = w & ~15u;
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(rr, (int)off, 0, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b128(rr, (int)(off + 16), 0, 0);
    const auto d = __builtin_amdgcn_raw_buffer_load_b128(rr, (int)(off + 32), 0, 0);
    wn = __builtin_amdgcn_raw_buffer_load_b32(rl, (int)(4u * (e0 + 64 + lane)), 0, 0);
    // the gathers and the next entries are in flight before any use
    __builtin_amdgcn_sched_barrier(0);
    const double pj0 = __hiloint2double((int)a[1], (int)a[0]), pj1 = __hiloint2double((int)a[3], (int)a[2]);
    const double pj2 = __hiloint2double((int)b[1], (int)b[0]), xj0 = __hiloint2double((int)b[3], (int)b[2]);
    const double xj1 = __hiloint2double((int)d[1], (int)d[0]), xj2 = __hiloint2double((int)d[3], (int)d[2]);
#pragma unroll
    for (int t = 0; t < kClAtoms; ++t) cl_term((w >> t) & 1u, pi[t], vi[t], pj0, pj1, pj2, xj0, xj1, xj2, v + 3 * t);
  }
  const double sum = cl_reduce(v, lane, red);
  if (lane >= 3 * kClAtoms) return;
  double out = sum + plus_i * xv;
  if (c0 != 0.0) out -= c0 * xv;
  out *= s1;
  if (rz) out -= c2 * zv;
  if (rout) rout[o] = out;
  if (yout) yout[12 * c + lane] = out;
  if (acc) acc[12 * c + lane] = av + mu * out;
}

// Records from a vector in the matrix's order: rec's vector part of slot q = v of atom
// cluster_atoms[q] (0 for a padding slot), or v of slot q itself when cluster_atoms is null
// (a vector already in slot order); acc (if given) = mu0 times the same, 3 per slot.
__global__ void k_rec_pack(int64_t slots, const int32_t *__restrict__ cluster_atoms, const double *__restrict__ v,
                           double *__restrict__ rec, double *__restrict__ acc, double mu0) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 3 * slots) return;
  const int64_t q = t / 3;
  const int k = (int)(t % 3);
  const int32_t at = cluster_atoms ? cluster_atoms[q] : (int32_t)q;
  const double x = at >= 0 ? v[3 * (int64_t)at + k] : 0.0;
  rec[6 * q + 3 + k] = x;
  if (acc) acc[t] = mu0 * x;
}

// u of atom cluster_atoms[q] = acc of slot q (padding slots skipped).
__global__ void k_rec_unpack(int64_t slots, const int32_t *__restrict__ cluster_atoms, const double *__restrict__ acc,
                             double *__restrict__ u) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 3 * slots) return;
  const int32_t at = cluster_atoms[t / 3];
  if (at >= 0) u[3 * (int64_t)at + t % 3] = acc[t];
}

// Several right-hand sides on the cluster layout: Y = (H + plus_i I) X, X and Y row-major
// [3n, nrhs]. One wave per (cluster, chunk of 4 columns); lane = 4 * slot + column, the 16
// slots take 16 list entries per iteration, so the 4 lanes of a slot share the neighbour's
// coordinates and read 4 adjacent columns of its vector rows. The halving reduction over the
// slot bits (lane bits 2..5) leaves lane l with value (l >> 2) of column (l & 3).
template <bool XCD>
__global__ void __launch_bounds__(64) k_matfree_clm(int64_t n, const double *__restrict__ xyz, int64_t ncl,
                                                     const int32_t *__restrict__ cluster_atoms,
                                                     const int64_t *__restrict__ cl_ptr,
                                                     const int32_t *__restrict__ base,
                                                     const uint32_t *__restrict__ entries, double plus_i,
                                                     const double *__restrict__ X, double *__restrict__ Y,
                                                     int nrhs) {
  const int64_t c = xcd_block(XCD);
  if (c >= ncl) return;
  const int lane = threadIdx.x;
  const int r = lane & 3, slot = lane >> 2;
  const int col = (int)blockIdx.y * 4 + r;
  const int colc = col < nrhs ? col : nrhs - 1;  // columns past nrhs compute on the last one, write nothing
  double pi[kClAtoms][3], vi[kClAtoms][3];
  double v[16];
#pragma unroll
  for (int t = 0; t < kClAtoms; ++t) {
    const int32_t at = cluster_atoms[kClAtoms * c + t];
    const int64_t i = at >= 0 ? at : 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      pi[t][k] = xyz[3 * i + k];
      vi[t][k] = X[(3 * i + k) * nrhs + colc];
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = 0.0;
  const int64_t estart = cl_ptr[c], e1 = cl_ptr[c + 1];
  const int64_t cb = base[c];
  const int64_t elast = e1 > estart ? e1 - 1 : estart;
  for (int64_t e0 = estart; e0 < e1; e0 += 16) {
    const int64_t e = e0 + slot;
    const uint32_t wl = entries[min(e, elast)];
    const uint32_t w = e < e1 ? wl : 0u;  // mask 0: no term
    const int64_t j = cb + (int64_t)(w >> 8);
    double pj[3], xj[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      pj[k] = xyz[3 * j + k];
      xj[k] = X[(3 * j + k) * nrhs + colc];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < kClAtoms; ++t) {
      const double d0 = pj[0] - pi[t][0], d1 = pj[1] - pi[t][1], d2 = pj[2] - pi[t][2];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const double w0 = vi[t][0] - xj[0], w1 = vi[t][1] - xj[1], w2 = vi[t][2] - xj[2];
      const bool on = ((w >> t) & 1u) && r2 > 0.0;
      const double s = on ? (d0 * w0 + d1 * w1 + d2 * w2) * recip_r2(r2) : 0.0;
      v[3 * t] += d0 * s;
      v[3 * t + 1] += d1 * s;
      v[3 * t + 2] += d2 * s;
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = swap_sum<32>(v[q], v[q + 8]);
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = swap_sum<16>(v[q], v[q + 4]);
#pragma unroll
  for (int o = 8, m = 2; o >= 4; o >>= 1, m >>= 1) {
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int q = 0; q < m; ++q) {
      const double lo = v[q] + __shfl_xor(v[q], o, 64);
      const double hi = v[q + m] + __shfl_xor(v[q + m], o, 64);
      v[q] = up ? hi : lo;
    }
  }
  const int m = lane >> 2;
  if (m >= 3 * kClAtoms || col >= nrhs) return;
  const int32_t at = cluster_atoms[kClAtoms * c + m / 3];
  if (at < 0) return;
  const int64_t dof = 3 * (int64_t)at + m % 3;
  Y[dof * nrhs + col] = v[0] + plus_i * X[dof * nrhs + col];
}

// Periodic images in x and y: the entry (i, j) carries a 9-bit mask of the images
// s = (sx, sy) of atom j within the cutoff (bit (sx + 1) * 3 + (sy + 1)), and the product sums
// d_s (d_s . (x_i - x_j)) / r_s^2 over them, d_s = (p_j + s box) - p_i. 16 lanes per row.
// Several right-hand sides (row-major [3n, nrhs]) when nrhs > 1; the epilogue applies to
// nrhs == 1 only.
template <typename VT>
__global__ void __launch_bounds__(256) k_matfree_pbc(int64_t n, const double *__restrict__ xyz,
                                                      const int64_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ col,
                                                      const uint16_t *__restrict__ images, double bx, double by,
                                                      double plus_i, const VT *__restrict__ x, VT *__restrict__ y,
                                                      int nrhs, double s1, double c0, const VT *__restrict__ z,
                                                      double c2, VT *__restrict__ acc, double mu,
                                                      const double *__restrict__ Q, const double *__restrict__ qcoef,
                                                      int mq) {
  constexpr int G = 16;
  const int lane = threadIdx.x & (G - 1);
  const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
  const int r = blockIdx.y;
  if (row >= n || r >= nrhs) return;
  const double pi0 = xyz[3 * row], pi1 = xyz[3 * row + 1], pi2 = xyz[3 * row + 2];
  const double vi0 = (double)x[(3 * row) * nrhs + r], vi1 = (double)x[(3 * row + 1) * nrhs + r],
               vi2 = (double)x[(3 * row + 2) * nrhs + r];
  double a0 = 0, a1 = 0, a2 = 0;
  for (int64_t b = rowptr[row] + lane; b < rowptr[row + 1]; b += G) {
    const int64_t j = col[b];
    const uint32_t m = images[b];
    const double w0 = vi0 - (double)x[(3 * j) * nrhs + r], w1 = vi1 - (double)x[(3 * j + 1) * nrhs + r],
                 w2 = vi2 - (double)x[(3 * j + 2) * nrhs + r];
    const double d2 = xyz[3 * j + 2] - pi2;
    for (int bit = 0; bit < 9; ++bit) {
      if (!((m >> bit) & 1u)) continue;
      const double d0 = (xyz[3 * j] + (double)(bit / 3 - 1) * bx) - pi0;
      const double d1 = (xyz[3 * j + 1] + (double)(bit % 3 - 1) * by) - pi1;
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      const double t = r2 > 0.0 ? (d0 * w0 + d1 * w1 + d2 * w2) * recip_r2(r2) : 0.0;
      a0 += d0 * t;
      a1 += d1 * t;
      a2 += d2 * t;
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off, G);
    a1 += __shfl_xor(a1, off, G);
    a2 += __shfl_xor(a2, off, G);
  }
  if (lane < 3) {
    const int64_t dof = 3 * row + lane;
    const double vi = lane == 0 ? vi0 : (lane == 1 ? vi1 : vi2);
    double out = (lane == 0 ? a0 : (lane == 1 ? a1 : a2)) + plus_i * vi;
    if (nrhs > 1) {
      y[dof * nrhs + r] = (VT)out;
      return;
    }
    if (c0 != 0.0) out -= c0 * vi;
    out *= s1;
    if (z) out -= c2 * (double)z[dof];
    if (qcoef) {
      double q = 0.0;
      const int64_t nd = 3 * n;
      for (int mm = 0; mm < mq; ++mm) q += qcoef[mm] * Q[mm * nd + dof];
      out += q;
    }
    y[dof] = (VT)out;
    if (acc) acc[dof] = (VT)((double)acc[dof] + mu * out);
  }
}

// Streaming copy used to measure the achievable HBM bandwidth (16 B per lane access).
// One 16-B element per lane and one pass over the grid (no grid-stride loop): on MI355X this
// form streams at ~6.2 TB/s, against ~4.6 TB/s for a grid-stride loop of the same accesses.
__global__ void __launch_bounds__(256) k_stream_copy(int64_t n4, const float4 *__restrict__ src,
                                                      float4 *__restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n4) dst[i] = src[i];
}

// ---------------------------------------------------------------- dense FP64 ---

template <int JT>
__global__ void __launch_bounds__(256) k_multidot_part(int64_t len, int J, const double *__restrict__ V, int64_t ld,
                                                        const double *__restrict__ u, double *__restrict__ part) {
  // part[blockIdx.x * J + j] for j in this block's tile [blockIdx.y*JT, +JT)
  const int j0 = blockIdx.y * JT;
  double a[JT];
#pragma unroll
  for (int t = 0; t < JT; ++t) a[t] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    const double ui = u[i];
#pragma unroll
    for (int t = 0; t < JT; ++t)
      if (j0 + t < J) a[t] += V[(int64_t)(j0 + t) * ld + i] * ui;
  }
  __shared__ double red[JT][4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < JT; ++t) {
    double v = a[t];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[t][w] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < JT && j0 + (int)threadIdx.x < J) {
    const int t = threadIdx.x;
    part[(int64_t)blockIdx.x * J + j0 + t] = ((red[t][0] + red[t][1]) + (red[t][2] + red[t][3]));
  }
}

// Final stage of the two-phase reductions: one 256-thread workgroup per output j sums
// the nparts partials part[p*J + j] with a fixed strided order and a fixed tree
// (deterministic, and no serial chain of nparts dependent loads).
__global__ void __launch_bounds__(256) k_reduce_parts(int nparts, int J, const double *__restrict__ part,
                                                      double *__restrict__ h, int do_sqrt) {
  const int j = blockIdx.x;
  if (j >= J) return;
  double s = 0.0;
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) s += part[(int64_t)p * J + j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (red[0] + red[1]) + (red[2] + red[3]);
    h[j] = do_sqrt ? sqrt(t) : t;
  }
}

__global__ void __launch_bounds__(256) k_gemv_update(int64_t len, int J, const double *__restrict__ V, int64_t ld,
                                                      const double *__restrict__ h, double scale, double *__restrict__ u,
                                                      double *__restrict__ part) {
  extern __shared__ double hs[];
  for (int t = threadIdx.x; t < J; t += blockDim.x) hs[t] = h[t] * scale;
  __syncthreads();
  double ss = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < J; ++j) s += V[(int64_t)j * ld + i] * hs[j];
    const double un = u[i] - s;
    u[i] = un;
    ss += un * un;
  }
  if (part) {
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
  }
}

__global__ void k_scale_div(int64_t len, const double *__restrict__ src, const double *__restrict__ div,
                            double *__restrict__ dst, double *__restrict__ dst2) {
  const double inv = 1.0 / div[0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = src[i] * inv;
    dst[i] = v;
    if (dst2) dst2[i] = v;
  }
}

__global__ void k_axpy_dev(int64_t len, const double *__restrict__ a, double sign, const double *__restrict__ x,
                           double *__restrict__ y) {
  const double s = a[0] * sign;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += s * x[i];
}

template <typename S, typename D>
__global__ void k_convert(int64_t len, const S *__restrict__ s, D *__restrict__ d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (D)s[i];
}

typedef double double4_t __attribute__((ext_vector_type(4)));

// out[q, i] = sum_r W[r, q] V[r, i] (q < k, r < m) with v_mfma_f64_16x16x4f64. A workgroup
// owns 16 NT outputs q0.. (NT <= 4 tiles of 16) and 64 NS columns; wave w takes the 16-column
// slices w, w + 4, ... of them, NS slices, each with NT accumulator tiles that share the
// slice's V operand. W goes through LDS in chunks of 64 rows, staged once per workgroup for all its
// NS x 64 columns. Operand layouts of the instruction: A (16 x 4): lane l holds
// A[l % 16][l / 16]; B (4 x 16): lane l holds B[l / 16][l % 16]; D (16 x 16): lane l, element
// e holds D[l / 16 + 4 e][l % 16]. Here A[q][r] = W[r][q] and B[r][i] = V[r][i]: each lane
// loads one element of V per slice and 4 rows (16 consecutive columns of 4 rows per wave).
template <int NS, int NT>
__global__ void __launch_bounds__(256) k_rr_combine_mfma(int64_t len, int m, int k, const double *__restrict__ V,
                                                          int64_t ld, const double *__restrict__ W,
                                                          double *__restrict__ out, int64_t ld_out) {
  __shared__ double ws[64][65];
  const int q0 = blockIdx.y * 16 * NT;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int64_t col[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) col[s] = (int64_t)blockIdx.x * (64 * NS) + 64 * s + 16 * w + (lane & 15);
  double4_t acc[NS][NT];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[s][t] = (double4_t){0.0, 0.0, 0.0, 0.0};
  for (int rc = 0; rc < m; rc += 64) {
    const int nr = (m - rc) < 64 ? (m - rc) : 64;
    __syncthreads();
    for (int t = tid; t < 64 * 16 * NT; t += 256) {
      const int r = t / (16 * NT), q = t % (16 * NT);
      ws[r][q] = (r < nr && q0 + q < k) ? W[(int64_t)(rc + r) * k + q0 + q] : 0.0;
    }
    __syncthreads();
    for (int r0 = 0; r0 < nr; r0 += 4) {
      const int r = r0 + (lane >> 4);
      double bv[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) bv[s] = (col[s] < len && r < nr) ? V[(int64_t)(rc + r) * ld + col[s]] : 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double av = ws[r0 + (lane >> 4)][16 * t + (lane & 15)];
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[s][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[s], acc[s][t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (col[s] >= len) continue;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int q = q0 + 16 * t + (lane >> 4) + 4 * e;
        if (q < k) out[(int64_t)q * ld_out + col[s]] = acc[s][t][e];
      }
  }
}

// Triangular parts of the symmetric block matrix as the reference's half storage
// sees it: part 0 = full A, 1 = L (scalar lower triangle incl. diagonal),
// 2 = L^T (scalar upper triangle incl. diagonal). One thread per scalar row.
template <typename MT, typename VT>
__global__ void k_bsr_part(int64_t n, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                           const MT *__restrict__ vals, const VT *__restrict__ x, VT *__restrict__ y, int part) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= 3 * n) return;
  const int64_t i = r / 3;
  const int p = (int)(r % 3);
  double acc = 0.0;
  for (int64_t b = rowptr[i]; b < rowptr[i + 1]; ++b) {
    const int64_t j = col[b];
    if (part == 1 && j > i) continue;
    if (part == 2 && j < i) continue;
    for (int q = 0; q < 3; ++q) {
      if (j == i && part == 1 && q > p) continue;
      if (j == i && part == 2 && q < p) continue;
      acc += (double)vals[9 * b + 3 * p + q] * (double)x[3 * j + q];
    }
  }
  y[r] = (VT)acc;
}

template <typename X, typename Y>
__global__ void k_axpby(int64_t len, double a, const X *__restrict__ x, double b, Y *__restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    if (b == 0.0)
      y[i] = (Y)(a * (double)x[i]);
    else
      y[i] = (Y)(a * (double)x[i] + b * (double)y[i]);
  }
}

// ---- delayed CGS2 (DCGS2) Lanczos step ------------------------------------------------
// Two reads of V[0..i] per step instead of CGS2's three: the second projection of the
// provisional vector ~v_i = V[i] (orthogonalised once, in step i - 1) is taken in step i's
// read together with the first projection of u = Op ~v_i, and one update pass then both
// corrects V[i] and orthogonalises u (Swirydowicz, Langou, Ananthan, Yang, Thomas 2020;
// Bielich, Langou, Thomas, Swirydowicz, Yamazaki, Boman 2022). The operator correction uses
// the projected matrix T: Op V[0..i-1] = V[0..i] T[0..i, 0..i-1].
constexpr int kDcgsMaxJ = 256;

// part[p*2J + j] = V_j . u (j < J), part[p*2J + J + j] = V_j . w (j < J), w = V[J-1]
template <int JT>
__global__ void __launch_bounds__(256) k_dual_dot_part(int64_t len, int J, const double *__restrict__ V, int64_t ld,
                                                        const double *__restrict__ u, const double *__restrict__ w,
                                                        double *__restrict__ part) {
  const int j0 = blockIdx.y * JT;
  double a[JT], b[JT];
#pragma unroll
  for (int t = 0; t < JT; ++t) a[t] = b[t] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    const double ui = u[i], wi = w[i];
#pragma unroll
    for (int t = 0; t < JT; ++t)
      if (j0 + t < J) {
        const double v = V[(int64_t)(j0 + t) * ld + i];
        a[t] += v * ui;
        b[t] += v * wi;
      }
  }
  __shared__ double red[2 * JT][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < JT; ++t) {
    double x = a[t], y = b[t];
    for (int off = 32; off > 0; off >>= 1) {
      x += __shfl_xor(x, off, 64);
      y += __shfl_xor(y, off, 64);
    }
    if (lane == 0) {
      red[t][wv] = x;
      red[JT + t][wv] = y;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < 2 * JT) {
    const int t = threadIdx.x % JT, which = threadIdx.x / JT;
    if (j0 + t < J)
      part[(int64_t)blockIdx.x * 2 * J + which * J + j0 + t] =
          ((red[threadIdx.x][0] + red[threadIdx.x][1]) + (red[threadIdx.x][2] + red[threadIdx.x][3]));
  }
}

__device__ __forceinline__ double block_sum256(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// One workgroup. sg = (s, g): s_j = V_j . u (j <= i), g_l = V_l . ~v_i (l < i), i = J - 1,
// u = A ~v_i for the provisional basis vector ~v_i. The corrected vector is
// v_i = ~v_i - sum_l g_l V_l (norm 1 + O(g^2)). Because A ~v_i - A v_i = sum_l g_l A V_l lies in
// span(V_0..V_i) for a Krylov basis, projecting u off that span gives A v_i's residual:
//   u' = u - sum_{j<i} s_j V_j - (v_i . u) v_i,  v_i . u = s_i - g . s,
// and alpha_i = v_i . A v_i = v_i . u - sum_l g_l (v_i . A V_l) = s_i - g . s - beta_{i-1} g_{i-1}
// (v_i . A V_l vanishes for l < i - 1, also for the Ritz vectors of a thick restart).
// coef[i] = v_i . u, coef[j] = s_j - coef[i] g_j (j < i), coef[J + l] = g_l.
__global__ void __launch_bounds__(256) k_dcgs2_coef(int J, const double *__restrict__ sg,
                                                     const double *__restrict__ beta, double *__restrict__ coef,
                                                     double *__restrict__ alpha_i) {
  __shared__ double red[4];
  const int i = J - 1;
  const double *s = sg, *g = sg + J;
  const int tid = threadIdx.x;
  double p = 0.0;
  for (int l = tid; l < i; l += blockDim.x) p += g[l] * s[l];
  const double ci = s[i] - block_sum256(p, red);
  for (int j = tid; j < i; j += blockDim.x) {
    coef[j] = s[j] - ci * g[j];
    coef[J + j] = g[j];
  }
  if (tid == 0) {
    coef[i] = ci;
    *alpha_i = i > 0 ? ci - beta[i - 1] * g[i - 1] : ci;
  }
}

// u -= sum_{j<=i} coef[j] V_j and V[i] -= sum_{l<i} coef[J+l] V_l in one read of V[0..i];
// part[block] = the block's sum of u'^2. LDS: coefficients staged in LDS (J <= 256);
// otherwise read from global memory (wave-uniform loads).
template <bool LDS>
__global__ void __launch_bounds__(256) k_dcgs2_update(int64_t len, int J, double *__restrict__ V, int64_t ld,
                                                       const double *__restrict__ coef, double *__restrict__ u,
                                                       double *__restrict__ part) {
  __shared__ double cs[LDS ? 2 * kDcgsMaxJ : 1];
  __shared__ double red[4];
  const int i = J - 1;
  const double *ca = coef, *cg = coef + J;
  if constexpr (LDS) {
    for (int t = threadIdx.x; t <= i; t += blockDim.x) cs[t] = coef[t];
    for (int t = threadIdx.x; t < i; t += blockDim.x) cs[kDcgsMaxJ + t] = coef[J + t];
    __syncthreads();
    ca = cs;
    cg = cs + kDcgsMaxJ;
  }
  double ss = 0.0;
  for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < len; d += (int64_t)gridDim.x * blockDim.x) {
    double su = 0.0, sv = 0.0;
    for (int j = 0; j < i; ++j) {
      const double v = V[(int64_t)j * ld + d];
      su += ca[j] * v;
      sv += cg[j] * v;
    }
    const double vi = V[(int64_t)i * ld + d];
    const double un = u[d] - (su + ca[i] * vi);
    u[d] = un;
    V[(int64_t)i * ld + d] = vi - sv;
    ss += un * un;
  }
  const double tot = block_sum256(ss, red);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

int grid_for(int64_t len, int bs, int cap) {
  int64_t g = (len + bs - 1) / bs;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

// rowbase[i] = first (smallest) column of row i, col16[b] = col[b] - rowbase[i]; flags a
// row whose span does not fit 16 bits (columns are ascending within a row).
__global__ void k_col16(int64_t n, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                        int32_t *__restrict__ rowbase, uint16_t *__restrict__ col16, int32_t *__restrict__ bad) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b0 = rowptr[i], b1 = rowptr[i + 1];
  const int32_t base = b1 > b0 ? col[b0] : 0;
  rowbase[i] = base;
  for (int64_t b = b0; b < b1; ++b) {
    const int64_t d = (int64_t)col[b] - base;
    if (d < 0 || d > 65535) {
      *bad = 1;
      col16[b] = 0;
    } else {
      col16[b] = (uint16_t)d;
    }
  }
}

template <typename MT, typename VT, typename CT, bool XCD = false>
inch_status launch_spmv(int64_t n, const int64_t *rowptr, const CT *col, const int32_t *rowbase, const void *vals,
                        const void *x, void *y, double s1, double c0, const void *z, double c2, void *acc, double mu,
                        const double *Q, const double *qcoef, int mq, int G, hipStream_t st) {
  const int bs = 256;
  const int64_t grid = (n * G + bs - 1) / bs;
#define INCH_SPMV_CASE(GG)                                                                                          \
  case GG:                                                                                                          \
    k_bsr_spmv<MT, VT, GG, CT, XCD><<<grid, bs, 0, st>>>(n, rowptr, col, rowbase, (const MT *)vals, (const VT *)x,      \
                                                    (VT *)y, (MT)s1,                                                \
                                                (MT)c0, (const VT *)z, (MT)c2, (VT *)acc, (MT)mu, Q, qcoef, mq); \
    break;
  switch (G) {
    INCH_SPMV_CASE(4)
    INCH_SPMV_CASE(8)
    INCH_SPMV_CASE(16)
    INCH_SPMV_CASE(32)
    INCH_SPMV_CASE(64)
    default: INCH_FAIL(INCH_ERR_INVALID, "inch_bsr_spmv: lanes_per_row must be 4/8/16/32/64 (got %d)", G);
  }
#undef INCH_SPMV_CASE
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

template <typename MT, typename VT>
inch_status launch_spmm(int64_t n, const int64_t *rowptr, const int32_t *col, const void *vals, const void *X,
                        void *Y, int nrhs, hipStream_t st) {
  constexpr int G = 8, R = 8;
  const int bs = 256;
  dim3 grid((unsigned)((n * G + bs - 1) / bs), (unsigned)((nrhs + R - 1) / R));
  k_bsr_spmm<MT, VT, G, R><<<grid, bs, 0, st>>>(n, rowptr, col, (const MT *)vals, (const VT *)X, (VT *)Y, nrhs);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

template <typename VT, int G, int U, typename CT, bool XCD>
void launch_matfree(int64_t n, const double *xyz, const int64_t *rowptr, const CT *col, const int32_t *rowbase,
                    double plus_i, const void *x, void *y, double s1, double c0, const void *z, double c2, void *acc,
                    double mu, const double *Q, const double *qcoef, int m_q, hipStream_t st) {
  const int bs = 256;
  const int64_t grid = (n * G + bs - 1) / bs;
  k_matfree_epi<VT, G, U, CT, XCD><<<grid, bs, 0, st>>>(n, xyz, rowptr, col, rowbase, plus_i, (const VT *)x, (VT *)y, s1,
                                                   c0, (const VT *)z, c2, (VT *)acc, mu, Q, qcoef, m_q);
}

template <typename VT, typename CT>
bool dispatch_matfree(int code, int64_t n, const double *xyz, const int64_t *rowptr, const CT *col,
                      const int32_t *rowbase, double plus_i, const void *x, void *y, double s1, double c0,
                      const void *z, double c2, void *acc, double mu, const double *Q, const double *qcoef, int m_q,
                      hipStream_t st) {
#define INCH_MF(GG, UU)                                                                                        \
  case GG * 10 + UU:                                                                                           \
    launch_matfree<VT, GG, UU, CT, false>(n, xyz, rowptr, col, rowbase, plus_i, x, y, s1, c0, z, c2, acc, mu, Q, \
                                          qcoef, m_q, st);                                                     \
    return true;                                                                                               \
  case 2000 + GG * 10 + UU:                                                                                    \
    launch_matfree<VT, GG, UU, CT, true>(n, xyz, rowptr, col, rowbase, plus_i, x, y, s1, c0, z, c2, acc, mu, Q,  \
                                         qcoef, m_q, st);                                                      \
    return true;
  switch (code) {
    INCH_MF(4, 1)
    INCH_MF(4, 2)
    INCH_MF(8, 1)
    INCH_MF(8, 2)
    INCH_MF(16, 1)
    INCH_MF(16, 2)
    INCH_MF(32, 1)
    INCH_MF(32, 2)
    default:
      return false;
  }
#undef INCH_MF
}

template <typename VT, int U, bool XCD>
void launch_matfree_cl(int64_t n, const double *xyz, int64_t ncl, const int32_t *catoms, const int64_t *cl_ptr,
                       const int32_t *base, const uint32_t *entries, double plus_i, const void *x, void *y, double s1,
                       double c0, const void *z, double c2, void *acc, double mu, const double *Q,
                       const double *qcoef, int m_q, hipStream_t st) {
  k_matfree_cl<VT, U, XCD><<<(unsigned)ncl, 64, 0, st>>>(n, xyz, ncl, catoms, cl_ptr, base, entries, plus_i,
                                                         (const VT *)x, (VT *)y, s1, c0, (const VT *)z, c2, (VT *)acc,
                                                         mu, Q, qcoef, m_q);
}

template <typename VT, bool XCD>
void launch_matfree_clp(int cpw, int64_t n, const double *xyz, int64_t ncl, const int32_t *catoms,
                        const int64_t *cl_ptr, const int32_t *base, const uint32_t *entries, double plus_i,
                        const void *x, void *y, double s1, double c0, const void *z, double c2, void *acc, double mu,
                        const double *Q, const double *qcoef, int m_q, hipStream_t st) {
  k_matfree_clp<VT, XCD><<<(unsigned)((ncl + cpw - 1) / cpw), 64, 0, st>>>(
      n, xyz, ncl, catoms, cl_ptr, base, entries, plus_i, (const VT *)x, (VT *)y, s1, c0, (const VT *)z, c2, (VT *)acc,
      mu, Q, qcoef, m_q, cpw);
}

template <typename VT>
bool dispatch_matfree_cl(int code, int64_t n, const double *xyz, int64_t ncl, const int32_t *catoms,
                         const int64_t *cl_ptr, const int32_t *base, const uint32_t *entries, double plus_i,
                         const void *x, void *y, double s1, double c0, const void *z, double c2, void *acc, double mu,
                         const double *Q, const double *qcoef, int m_q, hipStream_t st) {
#define INCH_CL(UU)                                                                                              \
  case UU:                                                                                                       \
    launch_matfree_cl<VT, UU, false>(n, xyz, ncl, catoms, cl_ptr, base, entries, plus_i, x, y, s1, c0, z, c2, acc, \
                                     mu, Q, qcoef, m_q, st);                                                     \
    return true;                                                                                                 \
  case 2000 + UU:                                                                                                \
    launch_matfree_cl<VT, UU, true>(n, xyz, ncl, catoms, cl_ptr, base, entries, plus_i, x, y, s1, c0, z, c2, acc,  \
                                    mu, Q, qcoef, m_q, st);                                                      \
    return true;
  switch (code) {
    INCH_CL(1)
    INCH_CL(2)
    INCH_CL(4)
    default:
      break;
  }
#undef INCH_CL
  // 100 + cpw: the persistent form, cpw clusters per wave
  const int shape = code % 2000, cpw = shape - 100;
  if (code < 4000 && (cpw == 4 || cpw == 8 || cpw == 16)) {
    if (code >= 2000)
      launch_matfree_clp<VT, true>(cpw, n, xyz, ncl, catoms, cl_ptr, base, entries, plus_i, x, y, s1, c0, z, c2, acc,
                                   mu, Q, qcoef, m_q, st);
    else
      launch_matfree_clp<VT, false>(cpw, n, xyz, ncl, catoms, cl_ptr, base, entries, plus_i, x, y, s1, c0, z, c2, acc,
                                    mu, Q, qcoef, m_q, st);
    return true;
  }
  return false;
}

}  // namespace

extern "C" {

inch_status inch_bsr_spmv(int64_t n, const int64_t *rowptr, const int32_t *col, const void *vals, int32_t mdt,
                          int32_t vdt, const void *x, void *y, double s1, double c0, const void *z, double c2,
                          void *acc, double mu, const double *Q, const double *qcoef, int32_t m_q,
                          int32_t lanes_per_row, void *stream) {
  INCH_REQUIRE(n > 0 && rowptr && col && vals && x && y, "inch_bsr_spmv: bad args");
  INCH_REQUIRE(!qcoef || (Q && m_q > 0), "inch_bsr_spmv: qcoef needs Q and m_q > 0");
  INCH_REQUIRE(x != y, "inch_bsr_spmv: x and y must not alias");
  hipStream_t st = (hipStream_t)stream;
  const int G = lanes_per_row > 0 ? lanes_per_row : 16;
  const int32_t *nb = nullptr;
  if (mdt == INCH_F32 && vdt == INCH_F32)
    return launch_spmv<float, float>(n, rowptr, col, nb, vals, x, y, s1, c0, z, c2, acc, mu, Q, qcoef, m_q, G, st);
  if (mdt == INCH_F32 && vdt == INCH_F64)
    return launch_spmv<float, double>(n, rowptr, col, nb, vals, x, y, s1, c0, z, c2, acc, mu, Q, qcoef, m_q, G, st);
  if (mdt == INCH_F64 && vdt == INCH_F64)
    return launch_spmv<double, double>(n, rowptr, col, nb, vals, x, y, s1, c0, z, c2, acc, mu, Q, qcoef, m_q, G, st);
  INCH_FAIL(INCH_ERR_INVALID, "inch_bsr_spmv: unsupported dtype pair (%d,%d)", mdt, vdt);
}

inch_status inch_bsr_spmv16(int64_t n, const int64_t *rowptr, const int32_t *rowbase, const uint16_t *col16,
                            const void *vals, int32_t mdt, int32_t vdt, const void *x, void *y, double s1, double c0,
                            const void *z, double c2, void *acc, double mu, const double *Q, const double *qcoef,
                            int32_t m_q, int32_t lanes_per_row, void *stream) {
  INCH_REQUIRE(n > 0 && rowptr && rowbase && col16 && vals && x && y, "inch_bsr_spmv16: bad args");
  INCH_REQUIRE(!qcoef || (Q && m_q > 0), "inch_bsr_spmv16: qcoef needs Q and m_q > 0");
  INCH_REQUIRE(x != y, "inch_bsr_spmv16: x and y must not alias");
  hipStream_t st = (hipStream_t)stream;
  // lanes_per_row + 1000: XCD-contiguous row mapping
  const bool xcd = lanes_per_row >= 1000;
  const int G = lanes_per_row % 1000 > 0 ? lanes_per_row % 1000 : 16;
  if (mdt == INCH_F32 && vdt == INCH_F32)
    return xcd ? launch_spmv<float, float, uint16_t, true>(n, rowptr, col16, rowbase, vals, x, y, s1, c0, z, c2, acc,
                                                            mu, Q, qcoef, m_q, G, st)
               : launch_spmv<float, float>(n, rowptr, col16, rowbase, vals, x, y, s1, c0, z, c2, acc, mu, Q, qcoef,
                                           m_q, G, st);
  if (mdt == INCH_F32 && vdt == INCH_F64)
    return xcd ? launch_spmv<float, double, uint16_t, true>(n, rowptr, col16, rowbase, vals, x, y, s1, c0, z, c2,
                                                             acc, mu, Q, qcoef, m_q, G, st)
               : launch_spmv<float, double>(n, rowptr, col16, rowbase, vals, x, y, s1, c0, z, c2, acc, mu, Q, qcoef,
                                            m_q, G, st);
  if (mdt == INCH_F64 && vdt == INCH_F64)
    return xcd ? launch_spmv<double, double, uint16_t, true>(n, rowptr, col16, rowbase, vals, x, y, s1, c0, z, c2,
                                                              acc, mu, Q, qcoef, m_q, G, st)
               : launch_spmv<double, double>(n, rowptr, col16, rowbase, vals, x, y, s1, c0, z, c2, acc, mu, Q, qcoef,
                                             m_q, G, st);
  INCH_FAIL(INCH_ERR_INVALID, "inch_bsr_spmv16: unsupported dtype pair (%d,%d)", mdt, vdt);
}

inch_status inch_bsr_col16(int64_t n, const int64_t *rowptr, const int32_t *col, int32_t *rowbase,
                           uint16_t *col16, void *stream) {
  INCH_REQUIRE(n > 0 && rowptr && col && rowbase && col16, "inch_bsr_col16: bad args");
  hipStream_t st = (hipStream_t)stream;
  int32_t *flag = nullptr;
  HIP_TRY(hipMalloc(&flag, sizeof(int32_t)));
  hipError_t e = hipMemsetAsync(flag, 0, sizeof(int32_t), st);
  if (e == hipSuccess) {
    k_col16<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, rowptr, col, rowbase, col16, flag);
    e = hipGetLastError();
  }
  int32_t bad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, flag, sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(flag);
  HIP_TRY(e);
  if (bad) INCH_FAIL(INCH_ERR_CAPACITY, "inch_bsr_col16: a row spans 65,536 columns or more");
  return INCH_OK;
}

inch_status inch_bsr_spmm(int64_t n, const int64_t *rowptr, const int32_t *col, const void *vals, int32_t mdt,
                          int32_t vdt, const void *X, void *Y, int32_t nrhs, void *stream) {
  INCH_REQUIRE(n > 0 && rowptr && col && vals && X && Y && nrhs > 0 && X != Y, "inch_bsr_spmm: bad args");
  hipStream_t st = (hipStream_t)stream;
  if (mdt == INCH_F32 && vdt == INCH_F32) return launch_spmm<float, float>(n, rowptr, col, vals, X, Y, nrhs, st);
  if (mdt == INCH_F32 && vdt == INCH_F64) return launch_spmm<float, double>(n, rowptr, col, vals, X, Y, nrhs, st);
  if (mdt == INCH_F64 && vdt == INCH_F64) return launch_spmm<double, double>(n, rowptr, col, vals, X, Y, nrhs, st);
  INCH_FAIL(INCH_ERR_INVALID, "inch_bsr_spmm: unsupported dtype pair (%d,%d)", mdt, vdt);
}

inch_status inch_anm_matfree_apply(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                                   double plus_i, const double *x, double *y, int32_t nrhs, void *stream) {
  INCH_REQUIRE(n > 0 && xyz && rowptr && col && x && y && nrhs > 0 && x != y, "inch_anm_matfree_apply: bad args");
  INCH_REQUIRE(nrhs <= 65535, "inch_anm_matfree_apply: nrhs too large");
  hipStream_t st = (hipStream_t)stream;
  constexpr int G = 16, R = 8;
  const int bs = 256;
  dim3 grid((unsigned)((n * G + bs - 1) / bs), (unsigned)((nrhs + R - 1) / R));
  k_matfree<G, R><<<grid, bs, 0, st>>>(n, xyz, rowptr, col, plus_i, x, y, nrhs);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_anm_matfree_spmv(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                                  double plus_i, int32_t vdt, const void *x, void *y, double s1, double c0,
                                  const void *z, double c2, void *acc, double mu, const double *Q,
                                  const double *qcoef, int32_t m_q, int32_t config, void *stream) {
  INCH_REQUIRE(n > 0 && xyz && rowptr && col && x && y && x != y, "inch_anm_matfree_spmv: bad args");
  INCH_REQUIRE(!qcoef || (Q && m_q > 0), "inch_anm_matfree_spmv: qcoef needs Q");
  hipStream_t st = (hipStream_t)stream;
  const int code = config > 0 ? config : 161;
  const int32_t *nb = nullptr;
  bool ok;
  if (vdt == INCH_F64)
    ok = dispatch_matfree<double>(code, n, xyz, rowptr, col, nb, plus_i, x, y, s1, c0, z, c2, acc, mu, Q, qcoef, m_q,
                                  st);
  else if (vdt == INCH_F32)
    ok = dispatch_matfree<float>(code, n, xyz, rowptr, col, nb, plus_i, x, y, s1, c0, z, c2, acc, mu, Q, qcoef, m_q,
                                 st);
  else
    INCH_FAIL(INCH_ERR_INVALID, "inch_anm_matfree_spmv: bad dtype");
  if (!ok) INCH_FAIL(INCH_ERR_INVALID, "inch_anm_matfree_spmv: config %d is not lanes*10+unroll with lanes in "
                                       "{4,8,16,32} and unroll in {1,2}, plus 2000 for XCD-contiguous rows", code);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_anm_matfree_pbc(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                                 const uint16_t *images, double box_x, double box_y, double plus_i, int32_t vdt,
                                 const void *x, void *y, int32_t nrhs, double s1, double c0, const void *z,
                                 double c2, void *acc, double mu, const double *Q, const double *qcoef, int32_t m_q,
                                 void *stream) {
  INCH_REQUIRE(n > 0 && xyz && rowptr && col && images && x && y && x != y && nrhs > 0 && nrhs <= 65535,
               "inch_anm_matfree_pbc: bad args");
  INCH_REQUIRE(nrhs == 1 || (!z && !acc && !qcoef && s1 == 1.0 && c0 == 0.0),
               "inch_anm_matfree_pbc: the epilogue needs nrhs == 1");
  INCH_REQUIRE(!qcoef || (Q && m_q > 0), "inch_anm_matfree_pbc: qcoef needs Q");
  hipStream_t st = (hipStream_t)stream;
  const int bs = 256;
  dim3 grid((unsigned)((n * 16 + bs - 1) / bs), (unsigned)nrhs);
  if (vdt == INCH_F64)
    k_matfree_pbc<double><<<grid, bs, 0, st>>>(n, xyz, rowptr, col, images, box_x, box_y, plus_i, (const double *)x,
                                               (double *)y, nrhs, s1, c0, (const double *)z, c2, (double *)acc, mu, Q,
                                               qcoef, m_q);
  else if (vdt == INCH_F32)
    k_matfree_pbc<float><<<grid, bs, 0, st>>>(n, xyz, rowptr, col, images, box_x, box_y, plus_i, (const float *)x,
                                              (float *)y, nrhs, s1, c0, (const float *)z, c2, (float *)acc, mu, Q,
                                              qcoef, m_q);
  else
    INCH_FAIL(INCH_ERR_INVALID, "inch_anm_matfree_pbc: bad dtype");
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_anm_matfree_spmv16(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *rowbase,
                                    const uint16_t *col16, double plus_i, int32_t vdt, const void *x, void *y,
                                    double s1, double c0, const void *z, double c2, void *acc, double mu,
                                    const double *Q, const double *qcoef, int32_t m_q, int32_t config, void *stream) {
  INCH_REQUIRE(n > 0 && xyz && rowptr && rowbase && col16 && x && y && x != y, "inch_anm_matfree_spmv16: bad args");
  INCH_REQUIRE(!qcoef || (Q && m_q > 0), "inch_anm_matfree_spmv16: qcoef needs Q");
  hipStream_t st = (hipStream_t)stream;
  const int code = config > 0 ? config : 161;
  bool ok;
  if (vdt == INCH_F64)
    ok = dispatch_matfree<double>(code, n, xyz, rowptr, col16, rowbase, plus_i, x, y, s1, c0, z, c2, acc, mu, Q,
                                  qcoef, m_q, st);
  else if (vdt == INCH_F32)
    ok = dispatch_matfree<float>(code, n, xyz, rowptr, col16, rowbase, plus_i, x, y, s1, c0, z, c2, acc, mu, Q,
                                 qcoef, m_q, st);
  else
    INCH_FAIL(INCH_ERR_INVALID, "inch_anm_matfree_spmv16: bad dtype");
  if (!ok) INCH_FAIL(INCH_ERR_INVALID, "inch_anm_matfree_spmv16: config %d is not lanes*10+unroll with lanes in "
                                       "{4,8,16,32} and unroll in {1,2}, plus 2000 for XCD-contiguous rows", code);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_anm_matfree_cl(int64_t n, const double *xyz, int64_t ncl, const int32_t *cluster_atoms,
                                const int64_t *cl_ptr, const int32_t *base, const uint32_t *entries, double plus_i,
                                int32_t vdt, const void *x, void *y, double s1, double c0, const void *z, double c2,
                                void *acc, double mu, const double *Q, const double *qcoef, int32_t m_q,
                                int32_t config, void *stream) {
  INCH_REQUIRE(n > 0 && xyz && cluster_atoms && cl_ptr && base && x && y && x != y,
               "inch_anm_matfree_cl: bad args");
  INCH_REQUIRE(ncl == (n + kClAtoms - 1) / kClAtoms, "inch_anm_matfree_cl: ncl must be ceil(n / 4)");
  INCH_REQUIRE(24 * n <= (int64_t)UINT32_MAX, "inch_anm_matfree_cl: n too large for 32-bit offsets");
  INCH_REQUIRE(!qcoef || (Q && m_q > 0), "inch_anm_matfree_cl: qcoef needs Q");
  hipStream_t st = (hipStream_t)stream;
  const int code = config > 0 ? config : 2002;
  bool ok;
  if (vdt == INCH_F64)
    ok = dispatch_matfree_cl<double>(code, n, xyz, ncl, cluster_atoms, cl_ptr, base, entries, plus_i, x, y, s1, c0,
                                     z, c2, acc, mu, Q, qcoef, m_q, st);
  else if (vdt == INCH_F32)
    ok = dispatch_matfree_cl<float>(code, n, xyz, ncl, cluster_atoms, cl_ptr, base, entries, plus_i, x, y, s1, c0,
                                    z, c2, acc, mu, Q, qcoef, m_q, st);
  else
    INCH_FAIL(INCH_ERR_INVALID, "inch_anm_matfree_cl: bad dtype");
  if (!ok) INCH_FAIL(INCH_ERR_INVALID, "inch_anm_matfree_cl: config %d is not an unroll in {1,2,4} or 100 + clusters "
                                       "per wave in {4,8,16}, plus 2000 for XCD-contiguous clusters", code);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_anm_matfree_rec(int64_t ncl, const double *rec_in, const double *rec_z, double *rec_out, double *y,
                                 const int64_t *cl_ptr, const uint32_t *entries, double plus_i, double s1, double c0,
                                 double c2, double *acc, double mu, int32_t xcd, void *stream) {
  INCH_REQUIRE(ncl > 0 && rec_in && (rec_out || y) && cl_ptr && entries && rec_in != rec_out &&
                   (!rec_z || rec_z != rec_out),
               "inch_anm_matfree_rec: bad args");
  INCH_REQUIRE(192 * ncl <= (int64_t)UINT32_MAX, "inch_anm_matfree_rec: %lld clusters exceed the 32-bit offsets",
               (long long)ncl);
  hipStream_t st = (hipStream_t)stream;
  if (xcd)
    k_matfree_rec<true><<<(unsigned)ncl, 64, 0, st>>>(ncl, rec_in, rec_z, rec_out, y, cl_ptr, entries, plus_i, s1,
                                                      c0, c2, acc, mu);
  else
    k_matfree_rec<false><<<(unsigned)ncl, 64, 0, st>>>(ncl, rec_in, rec_z, rec_out, y, cl_ptr, entries, plus_i, s1,
                                                       c0, c2, acc, mu);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_rec_pack(int64_t ncl, const int32_t *cluster_atoms, const double *v, double *rec, double *acc,
                          double mu0, void *stream) {
  INCH_REQUIRE(ncl > 0 && v && rec, "inch_rec_pack: bad args");
  const int64_t m = 3 * 4 * ncl;
  k_rec_pack<<<(unsigned)((m + 255) / 256), 256, 0, (hipStream_t)stream>>>(4 * ncl, cluster_atoms, v, rec, acc, mu0);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_rec_unpack(int64_t ncl, const int32_t *cluster_atoms, const double *acc, double *u, void *stream) {
  INCH_REQUIRE(ncl > 0 && cluster_atoms && acc && u, "inch_rec_unpack: bad args");
  const int64_t m = 3 * 4 * ncl;
  k_rec_unpack<<<(unsigned)((m + 255) / 256), 256, 0, (hipStream_t)stream>>>(4 * ncl, cluster_atoms, acc, u);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_anm_matfree_cl_apply(int64_t n, const double *xyz, int64_t ncl, const int32_t *cluster_atoms,
                                      const int64_t *cl_ptr, const int32_t *base, const uint32_t *entries,
                                      double plus_i, const double *x, double *y, int32_t nrhs, void *stream) {
  INCH_REQUIRE(n > 0 && xyz && cluster_atoms && cl_ptr && base && x && y && x != y && nrhs > 0,
               "inch_anm_matfree_cl_apply: bad args");
  INCH_REQUIRE(ncl == (n + kClAtoms - 1) / kClAtoms, "inch_anm_matfree_cl_apply: ncl must be ceil(n / 4)");
  INCH_REQUIRE(nrhs <= 4 * 65535, "inch_anm_matfree_cl_apply: nrhs too large");
  hipStream_t st = (hipStream_t)stream;
  dim3 grid((unsigned)ncl, (unsigned)((nrhs + 3) / 4));
  k_matfree_clm<true><<<grid, 64, 0, st>>>(n, xyz, ncl, cluster_atoms, cl_ptr, base, entries, plus_i, x, y, nrhs);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

int64_t inch_dcgs2_work_size(int32_t J) { return (int64_t)kRedBlocks * (2 * (int64_t)J + 1) + 4 * (int64_t)J; }

}  // extern "C"

namespace {

// The kernels of one delayed-CGS2 step on the scratch `work` (inch_dcgs2_work_size(J) doubles:
// per-block partial dots, the reduced dots, the coefficients, the norm partials). Every word
// of `work` that a kernel reads was written earlier in the same step, so the scratch needs no
// clearing between steps.
inch_status dcgs2_launch(int64_t len, int32_t J, double *V, int64_t ld, double *u, double *alpha,
                         const double *beta, double *beta_out, double *vnext, double *work, hipStream_t st) {
  constexpr int JT = 8;
  const int gx = grid_for(len, 256, kRedBlocks);
  double *part = work;
  double *sg = part + (int64_t)gx * 2 * J;
  double *coef = sg + 2 * J;
  double *npart = coef + 2 * J;
  dim3 grid(gx, (J + JT - 1) / JT);
  k_dual_dot_part<JT><<<grid, 256, 0, st>>>(len, J, V, ld, u, V + (int64_t)(J - 1) * ld, part);
  HIP_TRY(hipGetLastError());
  k_reduce_parts<<<2 * J, 256, 0, st>>>(gx, 2 * J, part, sg, 0);
  HIP_TRY(hipGetLastError());
  k_dcgs2_coef<<<1, 256, 0, st>>>(J, sg, beta, coef, alpha + (J - 1));
  HIP_TRY(hipGetLastError());
  if (J <= kDcgsMaxJ)
    k_dcgs2_update<true><<<gx, 256, 0, st>>>(len, J, V, ld, coef, u, npart);
  else
    k_dcgs2_update<false><<<gx, 256, 0, st>>>(len, J, V, ld, coef, u, npart);
  HIP_TRY(hipGetLastError());
  k_reduce_parts<<<1, 256, 0, st>>>(gx, 1, npart, beta_out, 1);
  HIP_TRY(hipGetLastError());
  if (vnext) {
    k_scale_div<<<grid_for(len, 256, 4096), 256, 0, st>>>(len, u, beta_out, vnext, nullptr);
    HIP_TRY(hipGetLastError());
  }
  return INCH_OK;
}

// Static scratch of the allocation-free step: one copy per GPU, sized for J <= kNoAllocMaxJ
// (2.0 MiB), held from the library's load until the process ends. Host code never passes the
// array's name to a kernel: its device address comes from hipGetSymbolAddress, once per
// device (with that device current), and is kept per device.
constexpr int32_t kNoAllocMaxJ = 256;
constexpr int64_t kNoAllocWords = (int64_t)kRedBlocks * (2 * kNoAllocMaxJ + 1) + 4 * kNoAllocMaxJ;
static __device__ __align__(256) double scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

inch_status scratch_for_current_device(double **out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) INCH_FAIL(INCH_ERR_INVALID, "scratch: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      INCH_FAIL(INCH_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return INCH_OK;
}

}  // namespace

extern "C" {

inch_status inch_dcgs2_step(int64_t len, int32_t J, double *V, int64_t ld, double *u, double *alpha,
                            const double *beta, double *beta_out, double *vnext, double *work, void *stream) {
  INCH_REQUIRE(len > 0 && J >= 1 && V && u && alpha && beta_out && work && ld >= len, "inch_dcgs2_step: bad args");
  INCH_REQUIRE(J == 1 || beta, "inch_dcgs2_step: beta needed for J > 1");
  return dcgs2_launch(len, J, V, ld, u, alpha, beta, beta_out, vnext, work, (hipStream_t)stream);
}

inch_status inch_dcgs2_step_noalloc(int64_t len, int32_t J, double *V, int64_t ld, double *u, double *alpha,
                                    const double *beta, double *beta_out, double *vnext, void *stream) {
  INCH_REQUIRE(len > 0 && J >= 1 && V && u && alpha && beta_out && ld >= len, "inch_dcgs2_step_noalloc: bad args");
  INCH_REQUIRE(J == 1 || beta, "inch_dcgs2_step_noalloc: beta needed for J > 1");
  if (J > kNoAllocMaxJ)
    INCH_FAIL(INCH_ERR_CAPACITY, "inch_dcgs2_step_noalloc: J = %d basis vectors exceed the %d its static scratch holds",
              J, kNoAllocMaxJ);
  double *work = nullptr;
  const inch_status s = scratch_for_current_device(&work);
  if (s != INCH_OK) return s;
  return dcgs2_launch(len, J, V, ld, u, alpha, beta, beta_out, vnext, work, (hipStream_t)stream);
}

inch_status inch_stream_copy(int64_t bytes, const void *src, void *dst, void *stream) {
  INCH_REQUIRE(bytes > 0 && bytes % 16 == 0 && src && dst, "inch_stream_copy: bytes must be a positive multiple of 16");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n4 = bytes / 16;
  k_stream_copy<<<(unsigned)((n4 + 255) / 256), 256, 0, st>>>(n4, (const float4 *)src, (float4 *)dst);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

int64_t inch_multidot_work_size(int32_t J) { return (int64_t)kRedBlocks * (J > 0 ? J : 1); }

inch_status inch_multidot(int64_t len, int32_t J, const double *V, int64_t ld, const double *u, double *h,
                          double *work, void *stream) {
  INCH_REQUIRE(len > 0 && J > 0 && V && u && h && work && ld >= len, "inch_multidot: bad args");
  hipStream_t st = (hipStream_t)stream;
  constexpr int JT = 8;
  const int gx = grid_for(len, 256, kRedBlocks);
  dim3 grid(gx, (J + JT - 1) / JT);
  k_multidot_part<JT><<<grid, 256, 0, st>>>(len, J, V, ld, u, work);
  HIP_TRY(hipGetLastError());
  k_reduce_parts<<<J, 256, 0, st>>>(gx, J, work, h, 0);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_gemv_update(int64_t len, int32_t J, const double *V, int64_t ld, const double *h, double scale,
                             double *u, double *nrm2_out, double *work, void *stream) {
  INCH_REQUIRE(len > 0 && J >= 0 && u && (J == 0 || (V && h)) && ld >= len, "inch_gemv_update: bad args");
  INCH_REQUIRE(!nrm2_out || work, "inch_gemv_update: nrm2 needs work");
  INCH_REQUIRE(J <= 8192, "inch_gemv_update: J too large");
  hipStream_t st = (hipStream_t)stream;
  const int gx = grid_for(len, 256, kRedBlocks);
  k_gemv_update<<<gx, 256, (size_t)(J > 0 ? J : 1) * sizeof(double), st>>>(len, J, V, ld, h, scale, u,
                                                                           nrm2_out ? work : nullptr);
  HIP_TRY(hipGetLastError());
  if (nrm2_out) {
    k_reduce_parts<<<1, 256, 0, st>>>(gx, 1, work, nrm2_out, 1);  // one workgroup, one output
    HIP_TRY(hipGetLastError());
  }
  return INCH_OK;
}

inch_status inch_vec_scale_div(int64_t len, const double *src, const double *div, double *dst, double *dst2,
                               void *stream) {
  INCH_REQUIRE(len > 0 && src && div && dst, "inch_vec_scale_div: bad args");
  hipStream_t st = (hipStream_t)stream;
  k_scale_div<<<grid_for(len, 256, 4096), 256, 0, st>>>(len, src, div, dst, dst2);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_vec_axpy_dev(int64_t len, const double *a, double sign, const double *x, double *y, void *stream) {
  INCH_REQUIRE(len > 0 && a && x && y, "inch_vec_axpy_dev: bad args");
  hipStream_t st = (hipStream_t)stream;
  k_axpy_dev<<<grid_for(len, 256, 4096), 256, 0, st>>>(len, a, sign, x, y);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_vec_convert(int64_t len, const void *src, int32_t sdt, void *dst, int32_t ddt, void *stream) {
  INCH_REQUIRE(len > 0 && src && dst, "inch_vec_convert: bad args");
  hipStream_t st = (hipStream_t)stream;
  const int g = grid_for(len, 256, 4096);
  if (sdt == INCH_F64 && ddt == INCH_F32)
    k_convert<double, float><<<g, 256, 0, st>>>(len, (const double *)src, (float *)dst);
  else if (sdt == INCH_F32 && ddt == INCH_F64)
    k_convert<float, double><<<g, 256, 0, st>>>(len, (const float *)src, (double *)dst);
  else if (sdt == INCH_F64 && ddt == INCH_F64)
    k_convert<double, double><<<g, 256, 0, st>>>(len, (const double *)src, (double *)dst);
  else if (sdt == INCH_F32 && ddt == INCH_F32)
    k_convert<float, float><<<g, 256, 0, st>>>(len, (const float *)src, (float *)dst);
  else
    INCH_FAIL(INCH_ERR_INVALID, "inch_vec_convert: bad dtypes");
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_bsr_spmv_part(int64_t n, const int64_t *rowptr, const int32_t *col, const void *vals,
                               int32_t mdt, int32_t vdt, const void *x, void *y, int32_t part, void *stream) {
  INCH_REQUIRE(n > 0 && rowptr && col && vals && x && y && x != y && part >= 0 && part <= 2,
               "inch_bsr_spmv_part: bad args");
  hipStream_t st = (hipStream_t)stream;
  const int64_t g = (3 * n + 255) / 256;
  if (mdt == INCH_F64 && vdt == INCH_F64)
    k_bsr_part<double, double><<<g, 256, 0, st>>>(n, rowptr, col, (const double *)vals, (const double *)x,
                                                  (double *)y, part);
  else if (mdt == INCH_F32 && vdt == INCH_F32)
    k_bsr_part<float, float><<<g, 256, 0, st>>>(n, rowptr, col, (const float *)vals, (const float *)x, (float *)y,
                                                part);
  else if (mdt == INCH_F32 && vdt == INCH_F64)
    k_bsr_part<float, double><<<g, 256, 0, st>>>(n, rowptr, col, (const float *)vals, (const double *)x,
                                                 (double *)y, part);
  else
    INCH_FAIL(INCH_ERR_INVALID, "inch_bsr_spmv_part: unsupported dtype pair");
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_vec_axpby(int64_t len, double a, const void *x, int32_t xdt, double b, void *y, int32_t ydt,
                           void *stream) {
  INCH_REQUIRE(len > 0 && x && y, "inch_vec_axpby: bad args");
  hipStream_t st = (hipStream_t)stream;
  const int g = grid_for(len, 256, 4096);
  if (xdt == INCH_F64 && ydt == INCH_F64)
    k_axpby<double, double><<<g, 256, 0, st>>>(len, a, (const double *)x, b, (double *)y);
  else if (xdt == INCH_F32 && ydt == INCH_F32)
    k_axpby<float, float><<<g, 256, 0, st>>>(len, a, (const float *)x, b, (float *)y);
  else if (xdt == INCH_F64 && ydt == INCH_F32)
    k_axpby<double, float><<<g, 256, 0, st>>>(len, a, (const double *)x, b, (float *)y);
  else if (xdt == INCH_F32 && ydt == INCH_F64)
    k_axpby<float, double><<<g, 256, 0, st>>>(len, a, (const float *)x, b, (double *)y);
  else
    INCH_FAIL(INCH_ERR_INVALID, "inch_vec_axpby: bad dtypes");
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_rr_combine(int64_t len, int32_t m, int32_t k, const double *V, int64_t ld, const double *W,
                            double *out, int64_t ld_out, void *stream) {
  INCH_REQUIRE(len > 0 && m > 0 && k > 0 && V && W && out && ld >= len && ld_out >= len, "inch_rr_combine: bad args");
  hipStream_t st = (hipStream_t)stream;
  // 4 output tiles with 4 slices (256 columns) per workgroup, fewer tiles with 8 slices (measured
  // best: 8 slices with 4 tiles took 0.74 against 0.68 ms for m = 192, k = 64, 900k columns)
  INCH_REQUIRE(len < ((int64_t)1 << 38) && k <= 64 * 65535, "inch_rr_combine: len or k too large");
  if (k > 32) {
    dim3 grid((unsigned)((len + 255) / 256), (unsigned)((k + 63) / 64));
    k_rr_combine_mfma<4, 4><<<grid, 256, 0, st>>>(len, m, k, V, ld, W, out, ld_out);
  } else if (k > 16) {
    dim3 grid((unsigned)((len + 511) / 512), 1);
    k_rr_combine_mfma<8, 2><<<grid, 256, 0, st>>>(len, m, k, V, ld, W, out, ld_out);
  } else {
    dim3 grid((unsigned)((len + 511) / 512), 1);
    k_rr_combine_mfma<8, 1><<<grid, 256, 0, st>>>(len, m, k, V, ld, W, out, ld_out);
  }
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

}  // extern "C"
