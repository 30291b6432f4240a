// Coordinate-graph network kernels (learned warm start): dense layers with float16
// weights/activations and float32 accumulation, neighbour-mean aggregation over the ANM
// neighbour graph, and their backward passes. Deterministic: every reduction over atoms
// is two-phase with fixed orders; no atomics.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.h"

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) INCH_FAIL(INCH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

namespace {

constexpr int kMaxD = 96;        // layer widths handled by the kernels (LDS: W and a 64-row tile)
constexpr int kRowsPerBlock = 64;
constexpr int kPartBlocks = 256;

// Y[i, o] = act( sum_k X[i, k] W[k, o] + b[o] ) ; X, W, Y float16; Ypre float32 (optional)
// One workgroup = 64 rows; W (Din x Dout) staged in LDS as float32.
__global__ void __launch_bounds__(256) k_dense_fwd(int64_t N, int Din, int Dout, const __half *__restrict__ X,
                                                    const __half *__restrict__ W, const float *__restrict__ b,
                                                    int relu, float *__restrict__ Ypre, __half *__restrict__ Y,
                                                    float *__restrict__ Y32) {
  __shared__ float ws[kMaxD * kMaxD];
  __shared__ float xs[kRowsPerBlock][kMaxD + 1];
  for (int t = threadIdx.x; t < Din * Dout; t += blockDim.x) ws[t] = __half2float(W[t]);
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock;
  for (int t = threadIdx.x; t < kRowsPerBlock * Din; t += blockDim.x) {
    const int r = t / Din, c = t % Din;
    xs[r][c] = (r0 + r < N) ? __half2float(X[(r0 + r) * Din + c]) : 0.f;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kRowsPerBlock * Dout; t += blockDim.x) {
    const int r = t / Dout, o = t % Dout;
    if (r0 + r >= N) continue;
    float acc = b ? b[o] : 0.f;
    for (int kk = 0; kk < Din; ++kk) acc += xs[r][kk] * ws[kk * Dout + o];
    const int64_t idx = (r0 + r) * Dout + o;
    if (Ypre) Ypre[idx] = acc;
    const float v = (relu && acc < 0.f) ? 0.f : acc;
    if (Y) Y[idx] = __float2half(v);
    if (Y32) Y32[idx] = v;
  }
}

// dX[i, k] = sum_o dY[i, o] W[k, o]   (float32 out)
__global__ void __launch_bounds__(256) k_dense_bwd_x(int64_t N, int Din, int Dout, const __half *__restrict__ W,
                                                      const float *__restrict__ dY, float *__restrict__ dX) {
  __shared__ float ws[kMaxD * kMaxD];
  __shared__ float ds[kRowsPerBlock][kMaxD + 1];
  for (int t = threadIdx.x; t < Din * Dout; t += blockDim.x) ws[t] = __half2float(W[t]);
  const int64_t r0 = (int64_t)blockIdx.x * kRowsPerBlock;
  for (int t = threadIdx.x; t < kRowsPerBlock * Dout; t += blockDim.x) {
    const int r = t / Dout, c = t % Dout;
    ds[r][c] = (r0 + r < N) ? dY[(r0 + r) * Dout + c] : 0.f;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kRowsPerBlock * Din; t += blockDim.x) {
    const int r = t / Din, kk = t % Din;
    if (r0 + r >= N) continue;
    float acc = 0.f;
    for (int o = 0; o < Dout; ++o) acc += ds[r][o] * ws[kk * Dout + o];
    dX[(r0 + r) * Din + kk] = acc;
  }
}

// partial dW[k, o] = sum_{i in block range} X[i, k] dY[i, o]; partial db[o] = sum dY[i, o]
// part layout per block: [Din*Dout + Dout]
__global__ void __launch_bounds__(256) k_dense_bwd_w(int64_t N, int Din, int Dout, const __half *__restrict__ X,
                                                      const float *__restrict__ dY, float *__restrict__ part) {
  const int64_t per = (N + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = blockIdx.x * per;
  const int64_t i1 = (i0 + per < N) ? i0 + per : N;
  float *dst = part + (int64_t)blockIdx.x * (Din * Dout + Dout);
  for (int t = threadIdx.x; t < Din * Dout + Dout; t += blockDim.x) {
    float acc = 0.f;
    if (t < Din * Dout) {
      const int kk = t / Dout, o = t % Dout;
      for (int64_t i = i0; i < i1; ++i) acc += __half2float(X[i * Din + kk]) * dY[i * Dout + o];
    } else {
      const int o = t - Din * Dout;
      for (int64_t i = i0; i < i1; ++i) acc += dY[i * Dout + o];
    }
    dst[t] = acc;
  }
}

// sums the per-block partials [dW | db] in a fixed order and splits them
__global__ void k_sum_part_wb(int nparts, int nw, int nb, const float *__restrict__ part, float *__restrict__ dW,
                              float *__restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = nw + nb;
  if (j >= m) return;
  float s = 0.f;
  for (int p = 0; p < nparts; ++p) s += part[(int64_t)p * m + j];
  if (j < nw)
    dW[j] = s;
  else
    db[j - nw] = s;
}

// mode 0 (forward):  Y[i] = sum_{j in N(i), j != i} X[j] / (deg_i - 1)
// mode 1 (adjoint):  Y[i] = sum_{j in N(i), j != i} X[j] / (deg_j - 1)   (N symmetric)
template <typename T, typename O>
__global__ void __launch_bounds__(256) k_aggregate(int64_t N, int D, const int64_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ col, const T *__restrict__ X,
                                                    O *__restrict__ Y, int mode) {
  const int64_t i = blockIdx.x;
  if (i >= N) return;
  const int64_t b0 = rowptr[i], b1 = rowptr[i + 1];
  const float inv_i = (b1 - b0 > 1) ? 1.f / (float)(b1 - b0 - 1) : 0.f;
  for (int d = threadIdx.x; d < D; d += blockDim.x) {
    float acc = 0.f;
    for (int64_t b = b0; b < b1; ++b) {
      const int64_t j = col[b];
      if (j == i) continue;
      float w = 1.f;
      if (mode == 1) {
        const int64_t dj = rowptr[j + 1] - rowptr[j] - 1;
        w = dj > 0 ? 1.f / (float)dj : 0.f;
      }
      acc += w * (float)X[j * D + d];
    }
    if (mode == 0) acc *= inv_i;
    Y[i * D + d] = (O)acc;
  }
}

__global__ void k_relu_bwd(int64_t n, const float *__restrict__ pre, const float *__restrict__ dY,
                           float *__restrict__ dPre) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dPre[i] = pre[i] > 0.f ? dY[i] : 0.f;
}

}  // namespace

extern "C" {

int64_t inch_gnn_work_size(int32_t Din, int32_t Dout) { return (int64_t)kPartBlocks * (Din * Dout + Dout); }

inch_status inch_gnn_dense_fwd(int64_t N, int32_t Din, int32_t Dout, const void *X, const void *W, const float *b,
                               int32_t relu, float *Ypre, void *Y, float *Y32, void *stream) {
  INCH_REQUIRE(N > 0 && Din > 0 && Dout > 0 && Din <= kMaxD && Dout <= kMaxD && X && W,
               "inch_gnn_dense_fwd: bad args (widths <= 96)");
  INCH_REQUIRE((int64_t)Din * Dout <= kMaxD * kMaxD, "inch_gnn_dense_fwd: layer too large");
  hipStream_t st = (hipStream_t)stream;
  const int64_t g = (N + kRowsPerBlock - 1) / kRowsPerBlock;
  k_dense_fwd<<<(unsigned)g, 256, 0, st>>>(N, Din, Dout, (const __half *)X, (const __half *)W, b, relu, Ypre,
                                           (__half *)Y, Y32);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_gnn_dense_bwd(int64_t N, int32_t Din, int32_t Dout, const void *X, const void *W, const float *dY,
                               float *dX, float *dW, float *db, float *work, void *stream) {
  INCH_REQUIRE(N > 0 && Din > 0 && Dout > 0 && Din <= kMaxD && Dout <= kMaxD && X && W && dY && dW && db && work,
               "inch_gnn_dense_bwd: bad args");
  hipStream_t st = (hipStream_t)stream;
  if (dX) {
    const int64_t g = (N + kRowsPerBlock - 1) / kRowsPerBlock;
    k_dense_bwd_x<<<(unsigned)g, 256, 0, st>>>(N, Din, Dout, (const __half *)W, dY, dX);
    HIP_TRY(hipGetLastError());
  }
  int gx = (int)((N + 63) / 64);
  if (gx > kPartBlocks) gx = kPartBlocks;
  k_dense_bwd_w<<<gx, 256, 0, st>>>(N, Din, Dout, (const __half *)X, dY, work);
  HIP_TRY(hipGetLastError());
  const int m = Din * Dout + Dout;
  k_sum_part_wb<<<(m + 255) / 256, 256, 0, st>>>(gx, Din * Dout, Dout, work, dW, db);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_gnn_aggregate(int64_t N, int32_t D, const int64_t *rowptr, const int32_t *col, const void *X,
                               int32_t xdt, void *Y, int32_t ydt, int32_t mode, void *stream) {
  INCH_REQUIRE(N > 0 && D > 0 && rowptr && col && X && Y && X != Y && (mode == 0 || mode == 1),
               "inch_gnn_aggregate: bad args");
  hipStream_t st = (hipStream_t)stream;
  const int bs = D <= 64 ? 64 : 128;
  // dtype codes: 0 float32, 2 float16
  if (xdt == 2 && ydt == 2)
    k_aggregate<__half, __half><<<(unsigned)N, bs, 0, st>>>(N, D, rowptr, col, (const __half *)X, (__half *)Y, mode);
  else if (xdt == 0 && ydt == 0)
    k_aggregate<float, float><<<(unsigned)N, bs, 0, st>>>(N, D, rowptr, col, (const float *)X, (float *)Y, mode);
  else if (xdt == 2 && ydt == 0)
    k_aggregate<__half, float><<<(unsigned)N, bs, 0, st>>>(N, D, rowptr, col, (const __half *)X, (float *)Y, mode);
  else
    INCH_FAIL(INCH_ERR_INVALID, "inch_gnn_aggregate: dtype pair (%d,%d)", xdt, ydt);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_gnn_relu_bwd(int64_t n, const float *pre, const float *dY, float *dPre, void *stream) {
  INCH_REQUIRE(n > 0 && pre && dY && dPre, "inch_gnn_relu_bwd: bad args");
  hipStream_t st = (hipStream_t)stream;
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  k_relu_bwd<<<(unsigned)g, 256, 0, st>>>(n, pre, dY, dPre);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

}  // extern "C"
