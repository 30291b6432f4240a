// Dense FP64 block contractions for the Rayleigh-Ritz / Gram steps (gfx950):
//   inch_gram            G = X Y^T for row blocks X [k1, n], Y [k2, n]  (MFMA f64 16x16x4)
//   inch_residual_norms  r_j = || Y[j] - theta_j X[j] ||_2
// Both are two-phase deterministic reductions over n: a fixed grid writes per-block
// partials that one small kernel sums in a fixed order.
#include <hip/hip_runtime.h>

#include "common.h"

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) INCH_FAIL(INCH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;    // output tile 64 x 64 per workgroup (4 waves x 16 rows)
constexpr int kChunk = 32;   // columns of n staged in LDS per step
constexpr int kGramBlocks = 256;   // fixed grid of the residual-norm reduction
constexpr int kGramGrid = 1024;    // most column blocks of a Gram product (at least 1,024 columns each)

// One workgroup = 4 waves; wave w owns output rows [16w, 16w+16) of the tile and the 4
// column blocks of 16. Operands are staged through LDS with coalesced row loads; the
// contraction dimension is the column index i of X and Y.
// v_mfma_f64_16x16x4f64 operand layout: A (16x4): lane l holds A[l%16][l/16];
// B (4x16): lane l holds B[l/16][l%16]; D (16x16): lane l, element r holds
// D[(l/16) + 4*r][l%16] (checked on gfx950 with one-hot operands).
__global__ void __launch_bounds__(256) k_gram_mfma(int64_t len, int k1, const double *__restrict__ X, int64_t ldx,
                                                    int k2, const double *__restrict__ Y, int64_t ldy,
                                                    double *__restrict__ part) {
  __shared__ double xs[kTile][kChunk + 1];
  __shared__ double ys[kTile][kChunk + 1];
  const int a0 = blockIdx.y * kTile, b0 = blockIdx.z * kTile;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double4_t acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const int64_t per = (len + gridDim.x - 1) / gridDim.x;
  const int64_t i_beg = blockIdx.x * per;
  const int64_t i_end = (i_beg + per < len) ? i_beg + per : len;
  for (int64_t i0 = i_beg; i0 < i_end; i0 += kChunk) {
    __syncthreads();
    // stage: 64 rows x 32 columns of each operand; thread t loads column t%32 of rows t/32 + 8s
    for (int s = 0; s < kTile * kChunk / 256; ++s) {
      const int r = tid / kChunk + s * (256 / kChunk);
      const int c = tid % kChunk;
      const int64_t i = i0 + c;
      xs[r][c] = (a0 + r < k1 && i < i_end) ? X[(int64_t)(a0 + r) * ldx + i] : 0.0;
      ys[r][c] = (b0 + r < k2 && i < i_end) ? Y[(int64_t)(b0 + r) * ldy + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kChunk; kk += 4) {
      const double av = xs[16 * w + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double bv = ys[16 * t + (lane & 15)][kk + (lane >> 4)];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
      }
    }
  }
  // write the 64x64 partial tile of this block
  double *dst = part + (int64_t)blockIdx.x * k1 * k2;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int a = a0 + 16 * w + (lane >> 4) + 4 * r;
      const int b = b0 + 16 * t + (lane & 15);
      if (a < k1 && b < k2) dst[(int64_t)a * k2 + b] = acc[t][r];
    }
  }
}

// Sum of nparts partials per output: thread per output when there are many outputs,
// workgroup per output (fixed tree) when there are few. Both orders are fixed.
__global__ void k_sum_parts(int nparts, int64_t m, const double *__restrict__ part, double *__restrict__ out,
                            int do_sqrt) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= m) return;
  double s = 0.0;
  for (int p = 0; p < nparts; ++p) s += part[(int64_t)p * m + j];
  out[j] = do_sqrt ? sqrt(s) : s;
}

__global__ void __launch_bounds__(256) k_sum_parts_wg(int nparts, int64_t m, const double *__restrict__ part,
                                                      double *__restrict__ out, int do_sqrt) {
  const int64_t j = blockIdx.x;
  if (j >= m) return;
  double s = 0.0;
  for (int p = threadIdx.x; p < nparts; p += blockDim.x) s += part[(int64_t)p * m + j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (red[0] + red[1]) + (red[2] + red[3]);
    out[j] = do_sqrt ? sqrt(t) : t;
  }
}

// partial sums of (Y[j,i] - theta_j X[j,i])^2 ; grid (gx, k)
__global__ void __launch_bounds__(256) k_resnorm(int64_t len, const double *__restrict__ X, int64_t ldx,
                                                  const double *__restrict__ Y, int64_t ldy,
                                                  const double *__restrict__ theta, double *__restrict__ part,
                                                  int k) {
  const int j = blockIdx.y;
  const double th = theta[j];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    const double r = Y[(int64_t)j * ldy + i] - th * X[(int64_t)j * ldx + i];
    s += r * r;
  }
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[(int64_t)blockIdx.x * k + j] = (red[0] + red[1]) + (red[2] + red[3]);
}

}  // namespace

extern "C" {

int64_t inch_gram_work_size(int32_t k1, int32_t k2) { return (int64_t)kGramGrid * k1 * k2; }

inch_status inch_gram(int64_t len, int32_t k1, const double *X, int64_t ldx, int32_t k2, const double *Y,
                      int64_t ldy, double *G, double *work, void *stream) {
  INCH_REQUIRE(len > 0 && k1 > 0 && k2 > 0 && X && Y && G && work && ldx >= len && ldy >= len,
               "inch_gram: bad args");
  INCH_REQUIRE(k1 <= 4096 && k2 <= 4096, "inch_gram: k too large");
  hipStream_t st = (hipStream_t)stream;
  // enough column blocks to fill the chip when k1 and k2 give one output tile (900k columns:
  // 879 blocks instead of 220 with 4,096 columns each, 0.72 -> see the dense-kernel profile)
  int64_t gx = (len + 1023) / 1024;
  if (gx > kGramGrid) gx = kGramGrid;
  if (gx < 1) gx = 1;
  dim3 grid((unsigned)gx, (k1 + kTile - 1) / kTile, (k2 + kTile - 1) / kTile);
  k_gram_mfma<<<grid, 256, 0, st>>>(len, k1, X, ldx, k2, Y, ldy, work);
  HIP_TRY(hipGetLastError());
  const int64_t m = (int64_t)k1 * k2;
  if (m >= 1024)
    k_sum_parts<<<(unsigned)((m + 255) / 256), 256, 0, st>>>((int)gx, m, work, G, 0);
  else
    k_sum_parts_wg<<<(unsigned)m, 256, 0, st>>>((int)gx, m, work, G, 0);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_residual_norms(int64_t len, int32_t k, const double *X, int64_t ldx, const double *Y, int64_t ldy,
                                const double *theta, double *out, double *work, void *stream) {
  INCH_REQUIRE(len > 0 && k > 0 && X && Y && theta && out && work && ldx >= len && ldy >= len,
               "inch_residual_norms: bad args");
  INCH_REQUIRE(k <= 65535, "inch_residual_norms: k too large");
  hipStream_t st = (hipStream_t)stream;
  int64_t gx = (len + 255) / 256;
  if (gx > kGramBlocks) gx = kGramBlocks;
  dim3 grid((unsigned)gx, (unsigned)k);
  k_resnorm<<<grid, 256, 0, st>>>(len, X, ldx, Y, ldy, theta, work, k);
  HIP_TRY(hipGetLastError());
  k_sum_parts_wg<<<(unsigned)k, 256, 0, st>>>((int)gx, k, work, out, 1);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ CGS2 ----
// Fused classical Gram-Schmidt with re-orthogonalisation for one Lanczos step, basis
// V[:J] (row-major, J <= 256), given h1 = V[:J] u (pass-1 coefficients):
//   pass B:  u' = u - V^T h1  (in place),  h2 = V[:J] u'  and ||u'||^2  (one read of V)
//   pass C:  u'' = u' - V^T h2,  vnext = u'' / ||u'||                      (one read of V)
// ||u'|| is the Lanczos beta (taken before the second pass, as in the reference).
namespace {

constexpr int kCgsMaxJ = 256;
constexpr int kCgsBlocks = 2048;

// Pass B over column tiles of C columns (C * J <= 4096 doubles of LDS): stage V[:J, tile]
// in LDS, form u' for the tile (256/C lanes per column, fixed butterfly), then thread j
// accumulates V[j, tile] . u'[tile] for its row. Partials per workgroup are summed in a
// fixed order by k_sum_parts_wg.
template <int C>
__global__ void __launch_bounds__(256) k_cgs_pass_b(int64_t len, int J, const double *__restrict__ V, int64_t ld,
                                                     const double *__restrict__ h1, double *__restrict__ u,
                                                     double *__restrict__ part) {
  constexpr int LPC = 256 / C;        // lanes per column
  constexpr int RS = 256 / C;         // rows staged per pass of the workgroup
  constexpr int NR = (4096 / C) / RS; // staging loads per thread (J <= 4096 / C)
  __shared__ double vs[4096 / C][C + 1];
  __shared__ double hs[kCgsMaxJ];
  __shared__ double up[C];
  const int tid = threadIdx.x;
  for (int t = tid; t < J; t += 256) hs[t] = h1[t];
  double acc = 0.0, ss = 0.0;
  const int64_t ntiles = (len + C - 1) / C;
  // the next tile's V[:J, tile] is fetched into registers while the current one is used
  const int c_ld = tid % C, r_ld = tid / C;
  double pre[NR];
  auto fetch = [&](int64_t tile) {
    const int64_t i = tile * C + c_ld;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const int r = r_ld + q * RS;
      pre[q] = (r < J && i < len) ? V[(int64_t)r * ld + i] : 0.0;
    }
  };
  if (blockIdx.x < ntiles) fetch(blockIdx.x);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t c0 = tile * C;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const int r = r_ld + q * RS;
      if (r < J) vs[r][c_ld] = pre[q];
    }
    if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x);
    __syncthreads();
    {
      const int c = tid / LPC, pl = tid % LPC;
      double s = 0.0;
      for (int r = pl; r < J; r += LPC) s += vs[r][c] * hs[r];
#pragma unroll
      for (int off = LPC / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPC);
      if (pl == 0) {
        const int64_t i = c0 + c;
        const double un = (i < len) ? u[i] - s : 0.0;
        up[c] = un;
        if (i < len) u[i] = un;
      }
    }
    __syncthreads();
    if (tid < J) {
#pragma unroll
      for (int c = 0; c < C; ++c) acc += vs[tid][c] * up[c];
    }
    if (tid == 0) {
#pragma unroll
      for (int c = 0; c < C; ++c) ss += up[c] * up[c];
    }
  }
  double *dst = part + (int64_t)blockIdx.x * (J + 1);
  if (tid < J) dst[tid] = acc;
  if (tid == 0) dst[J] = ss;
}

__global__ void __launch_bounds__(256) k_cgs_pass_c(int64_t len, int J, const double *__restrict__ V, int64_t ld,
                                                     const double *__restrict__ h2, double *__restrict__ u,
                                                     double *__restrict__ vnext) {
  __shared__ double hs[kCgsMaxJ + 1];
  for (int t = threadIdx.x; t <= J; t += blockDim.x) hs[t] = h2[t];
  __syncthreads();
  const double inv = 1.0 / sqrt(hs[J]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < J; ++j) s += V[(int64_t)j * ld + i] * hs[j];
    const double un = u[i] - s;
    u[i] = un;
    if (vnext) vnext[i] = un * inv;
  }
}

__global__ void k_sqrt_copy(const double *__restrict__ src, double *__restrict__ dst) { dst[0] = sqrt(src[0]); }

}  // namespace

extern "C" {

int64_t inch_cgs2_work_size(int32_t J) { return (int64_t)kCgsBlocks * (J + 1) + (J + 1); }

inch_status inch_cgs2_reorth(int64_t len, int32_t J, const double *V, int64_t ld, const double *h1, double *u,
                             double *vnext, double *beta_out, double *work, void *stream) {
  INCH_REQUIRE(len > 0 && J > 0 && J <= kCgsMaxJ && V && h1 && u && beta_out && work && ld >= len,
               "inch_cgs2_reorth: bad args (J must be 1..256)");
  hipStream_t st = (hipStream_t)stream;
  const int C = J <= 64 ? 64 : (J <= 128 ? 32 : 16);
  const int64_t ntiles = (len + C - 1) / C;
  const int gx = (int)(ntiles < kCgsBlocks ? ntiles : kCgsBlocks);
  double *part = work;
  double *h2 = work + (int64_t)kCgsBlocks * (J + 1);
  if (C == 64)
    k_cgs_pass_b<64><<<gx, 256, 0, st>>>(len, J, V, ld, h1, u, part);
  else if (C == 32)
    k_cgs_pass_b<32><<<gx, 256, 0, st>>>(len, J, V, ld, h1, u, part);
  else
    k_cgs_pass_b<16><<<gx, 256, 0, st>>>(len, J, V, ld, h1, u, part);
  HIP_TRY(hipGetLastError());
  k_sum_parts_wg<<<(unsigned)(J + 1), 256, 0, st>>>(gx, (int64_t)(J + 1), part, h2, 0);
  HIP_TRY(hipGetLastError());
  int64_t gc = (len + 255) / 256;
  if (gc > 4096) gc = 4096;
  k_cgs_pass_c<<<(unsigned)gc, 256, 0, st>>>(len, J, V, ld, h2, u, vnext);
  HIP_TRY(hipGetLastError());
  k_sqrt_copy<<<1, 1, 0, st>>>(h2 + J, beta_out);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

}  // extern "C"
