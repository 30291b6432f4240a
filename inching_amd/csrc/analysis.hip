// Downstream analysis of computed modes (Burn/Visualisation/T2.py:43-160): per-mode
// square fluctuations weighted by the relative variance of each mode, and the
// eigenvalue-weighted cross-correlation maps between pairs of modes.
//
// Modes are float64 [k, N, 3] (mode, atom, xyz) in the caller's atom order. Every kernel
// streams the modes once (square fluctuations) or writes one N x N map per pair with the two
// modes' rows read from L2 (cross-correlation), so both are HBM-bound.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"

namespace {

// ratio_var[m] = r_m / max_m r_m, r_m = 1 / sqrt(|lambda_m|) (one block; k values).
__global__ void __launch_bounds__(256) k_ratio_var(int32_t k, const double *__restrict__ eigval,
                                                   double *__restrict__ ratio_var) {
  __shared__ double smax[256];
  double mx = 0.0;
  for (int m = threadIdx.x; m < k; m += blockDim.x) mx = fmax(mx, 1.0 / sqrt(fabs(eigval[m])));
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    __syncthreads();
  }
  const double top = smax[0];
  for (int m = threadIdx.x; m < k; m += blockDim.x) ratio_var[m] = (1.0 / sqrt(fabs(eigval[m]))) / top;
}

// One thread per atom: sfpm[m, a] = |v_m[a]|, ratio[m, a] = ratio_var[m] sfpm[m, a],
// sqfluct[a] = sum_m ratio[m, a] (modes summed in order).
__global__ void __launch_bounds__(256) k_sqfluct(int32_t k, int64_t N, const double *__restrict__ eigvec,
                                                 const double *__restrict__ ratio_var, double *__restrict__ sfpm,
                                                 double *__restrict__ ratio, double *__restrict__ sqfluct) {
  const int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (a >= N) return;
  double acc = 0.0;
  for (int m = 0; m < k; ++m) {
    const double *v = eigvec + ((int64_t)m * N + a) * 3;
    const double s = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double r = ratio_var[m] * s;
    sfpm[(int64_t)m * N + a] = s;
    ratio[(int64_t)m * N + a] = r;
    acc += r;
  }
  sqfluct[a] = acc;
}

// out[a, b] = L((v_i[a] . v_j[b]) lambda_i lambda_j) for one pair (i, j) of modes, with the
// reference's log-modulus transform L(s) = sign(s) log2(|s 100^precision| + 1) applied in
// float32 to the float32 value of the product (the reference forms the product in float64,
// multiplies by lambda_i, then by lambda_j, and accumulates it into a float32 tensor before
// the transform). A 64 x 64 tile per workgroup: the tile's 64 rows of v_i and 64 rows of v_j
// go through LDS; each thread writes 16 entries of one column of the tile, consecutive
// threads consecutive b (coalesced stores).
__global__ void __launch_bounds__(256) k_cross_corr(int64_t N, const double *__restrict__ vi_g,
                                                    const double *__restrict__ vj_g, double li, double lj,
                                                    float scale, float *__restrict__ out) {
  __shared__ double vi[64][3], vj[64][3];
  const int64_t a0 = (int64_t)blockIdx.y * 64, b0 = (int64_t)blockIdx.x * 64;
  const int t = threadIdx.x;
  if (t < 64) {
    const int64_t a = a0 + t;
    for (int c = 0; c < 3; ++c) vi[t][c] = a < N ? vi_g[3 * a + c] : 0.0;
  } else if (t < 128) {
    const int64_t b = b0 + (t - 64);
    for (int c = 0; c < 3; ++c) vj[t - 64][c] = b < N ? vj_g[3 * b + c] : 0.0;
  }
  __syncthreads();
  const int col = t & 63;
  const int64_t b = b0 + col;
  if (b >= N) return;
  for (int r = t >> 6; r < 64; r += 4) {
    const int64_t a = a0 + r;
    if (a >= N) break;
    const double d = vi[r][0] * vj[col][0] + vi[r][1] * vj[col][1] + vi[r][2] * vj[col][2];
    const float s = (float)((d * li) * lj) * scale;
    const float l = log2f(fabsf(s) + 1.0f);
    out[a * N + b] = s > 0.0f ? l : (s < 0.0f ? -l : 0.0f);
  }
}

}  // namespace

extern "C" inch_status inch_sqfluct(int32_t k, int64_t N, const double *eigval, const double *eigvec,
                                    double *ratio_var, double *sfpm, double *ratio, double *sqfluct, void *stream) {
  INCH_REQUIRE(k > 0 && N > 0 && eigval && eigvec && ratio_var && sfpm && ratio && sqfluct,
               "inch_sqfluct: bad args");
  hipStream_t st = (hipStream_t)stream;
  k_ratio_var<<<1, 256, 0, st>>>(k, eigval, ratio_var);
  k_sqfluct<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(k, N, eigvec, ratio_var, sfpm, ratio, sqfluct);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) INCH_FAIL(INCH_ERR_HIP, "inch_sqfluct: %s", hipGetErrorString(e));
  return INCH_OK;
}

extern "C" inch_status inch_cross_correlation(int32_t P, const int32_t *pairs, int32_t k, int64_t N,
                                              const double *eigval_host, const double *eigvec, int32_t precision,
                                              float *out, void *stream) {
  INCH_REQUIRE(P > 0 && k > 0 && N > 0 && pairs && eigval_host && eigvec && out && precision >= 0 && precision <= 4,
               "inch_cross_correlation: bad args");
  INCH_REQUIRE((N + 63) / 64 <= 65535, "inch_cross_correlation: too many atoms for the grid");
  for (int32_t p = 0; p < P; ++p)
    INCH_REQUIRE(pairs[2 * p] >= 0 && pairs[2 * p] < k && pairs[2 * p + 1] >= 0 && pairs[2 * p + 1] < k,
                 "inch_cross_correlation: pair %d = (%d, %d) outside the %d modes", p, pairs[2 * p], pairs[2 * p + 1], k);
  hipStream_t st = (hipStream_t)stream;
  float scale = 1.0f;
  for (int32_t q = 0; q < precision; ++q) scale *= 100.0f;   // 100^precision, exact in float32 up to 100^5
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((N + 63) / 64));
  for (int32_t p = 0; p < P; ++p) {
    const int32_t i = pairs[2 * p], j = pairs[2 * p + 1];
    k_cross_corr<<<grid, 256, 0, st>>>(N, eigvec + (int64_t)i * N * 3, eigvec + (int64_t)j * N * 3, eigval_host[i],
                                       eigval_host[j], scale, out + (int64_t)p * N * N);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) INCH_FAIL(INCH_ERR_HIP, "inch_cross_correlation: %s", hipGetErrorString(e));
  return INCH_OK;
}
