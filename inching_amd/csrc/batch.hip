// Segmented (batched) FP64 kernels: many independent structures stored as contiguous dof
// segments of one long vector. Work is cut into chunks that never straddle a segment;
// chunk c covers dofs [cbeg[c], cend[c]) of segment cseg[c], and the chunks of segment s
// are sptr[s] .. sptr[s+1]-1. Every reduction is two-phase with fixed orders
// (deterministic): per-chunk partials, then a per-segment sum over its chunks.
#include <hip/hip_runtime.h>

#include "common.h"

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) INCH_FAIL(INCH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

namespace {

constexpr int kSegMaxJ = 512;

__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// part[c*J + j] = sum_{i in chunk c} V[j*ld + i] * u[i]; blockIdx.y picks the tile of 8 rows
__global__ void __launch_bounds__(256) k_seg_dot_part(const int64_t *__restrict__ cbeg, const int64_t *__restrict__ cend,
                                                       int J, const double *__restrict__ V, int64_t ld,
                                                       const double *__restrict__ u, double *__restrict__ part) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int64_t b = cbeg[c], e = cend[c];
  {
    const int j0 = 8 * (int)blockIdx.y;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
      const double ui = u[i];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (j0 + t < J) a[t] += V[(int64_t)(j0 + t) * ld + i] * ui;
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const double s = block_sum(a[t], red);
      if (threadIdx.x == 0 && j0 + t < J) part[(int64_t)c * J + j0 + t] = s;
    }
  }
}

// h[s*J + j] = sum over the chunks of segment s (optionally sqrt)
__global__ void k_seg_reduce(int S, int J, const int64_t *__restrict__ sptr, const double *__restrict__ part,
                             double *__restrict__ h, int do_sqrt) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)S * J) return;
  const int s = (int)(t / J), j = (int)(t % J);
  double acc = 0.0;
  for (int64_t c = sptr[s]; c < sptr[s + 1]; ++c) acc += part[c * J + j];
  h[t] = do_sqrt ? sqrt(acc) : acc;
}

// u[i] -= scale * sum_j V[j*ld + i] * h[s*hld + j] for i in chunk c of segment s; part[c] = sum u_new^2
__global__ void __launch_bounds__(256) k_seg_update(const int64_t *__restrict__ cbeg, const int64_t *__restrict__ cend,
                                                     const int32_t *__restrict__ cseg, int J,
                                                     const double *__restrict__ V, int64_t ld,
                                                     const double *__restrict__ h, int hld, double scale,
                                                     double *__restrict__ u, double *__restrict__ part) {
  __shared__ double hs[kSegMaxJ];
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int s = cseg[c];
  for (int t = threadIdx.x; t < J; t += blockDim.x) hs[t] = h[(int64_t)s * hld + t] * scale;
  __syncthreads();
  double ss = 0.0;
  for (int64_t i = cbeg[c] + threadIdx.x; i < cend[c]; i += blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < J; ++j) acc += V[(int64_t)j * ld + i] * hs[j];
    const double un = u[i] - acc;
    u[i] = un;
    ss += un * un;
  }
  if (part) {
    const double tot = block_sum(ss, red);
    if (threadIdx.x == 0) part[c] = tot;
  }
}

// Rigid-body projection per segment from the coordinates instead of six stored basis rows.
// With r_i = p_i - cen[s] (atom i of segment s relative to the segment's centroid), the
// translations e_a and the rotations e_k x r_i of the segment are orthogonal to each other
// (sum_i r_i = 0), the translations have Gram matrix n_s I and the rotations
// K_s = sum_i (|r_i|^2 I - r_i r_i^T). So the projection of x onto their span is, at atom i,
// mean_s(x) + (K_s^{-1} sum_j r_j x x_j) x r_i: six moments per segment, then one pass.
// part[c*6 + a] = sum of component a of x over chunk c (a < 3), part[c*6 + 3 + a] = component
// a of sum r_i x x_i over the chunk.
__global__ void __launch_bounds__(256) k_seg_rigid_moments(const int64_t *__restrict__ cbeg,
                                                            const int64_t *__restrict__ cend,
                                                            const int32_t *__restrict__ cseg,
                                                            const double *__restrict__ xyz,
                                                            const double *__restrict__ cen,
                                                            const double *__restrict__ x, double *__restrict__ part) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int s = cseg[c];
  const double q0 = cen[3 * s], q1 = cen[3 * s + 1], q2 = cen[3 * s + 2];
  double m[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t d = cbeg[c] + threadIdx.x; d < cend[c]; d += blockDim.x) {
    const int64_t at = d / 3;
    const int a = (int)(d - 3 * at);
    const double v = x[d];
    const double r0 = xyz[3 * at] - q0, r1 = xyz[3 * at + 1] - q1, r2 = xyz[3 * at + 2] - q2;
    if (a == 0) {            // r x (v e_x) = v (0, r2, -r1)
      m[0] += v;
      m[4] += r2 * v;
      m[5] -= r1 * v;
    } else if (a == 1) {     // r x (v e_y) = v (-r2, 0, r0)
      m[1] += v;
      m[3] -= r2 * v;
      m[5] += r0 * v;
    } else {                 // r x (v e_z) = v (r1, -r0, 0)
      m[2] += v;
      m[3] += r1 * v;
      m[4] -= r0 * v;
    }
  }
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const double tot = block_sum(m[t], red);
    if (threadIdx.x == 0) part[(int64_t)c * 6 + t] = tot;
  }
}

// y -= scale * P x over chunk c of segment s: g[s*6..] = the segment's moments of x,
// kinv[s*9..] = K_s^{-1} (row-major), ninv[s] = 1 / n_s.
__global__ void __launch_bounds__(256) k_seg_rigid_update(const int64_t *__restrict__ cbeg,
                                                           const int64_t *__restrict__ cend,
                                                           const int32_t *__restrict__ cseg,
                                                           const double *__restrict__ xyz,
                                                           const double *__restrict__ cen,
                                                           const double *__restrict__ kinv,
                                                           const double *__restrict__ ninv,
                                                           const double *__restrict__ g, double scale,
                                                           double *__restrict__ y) {
  const int c = blockIdx.x;
  const int s = cseg[c];
  const double *gs = g + 6 * (int64_t)s, *K = kinv + 9 * (int64_t)s;
  const double t0 = gs[0] * ninv[s], t1 = gs[1] * ninv[s], t2 = gs[2] * ninv[s];
  const double f0 = K[0] * gs[3] + K[1] * gs[4] + K[2] * gs[5];
  const double f1 = K[3] * gs[3] + K[4] * gs[4] + K[5] * gs[5];
  const double f2 = K[6] * gs[3] + K[7] * gs[4] + K[8] * gs[5];
  const double q0 = cen[3 * s], q1 = cen[3 * s + 1], q2 = cen[3 * s + 2];
  for (int64_t d = cbeg[c] + threadIdx.x; d < cend[c]; d += blockDim.x) {
    const int64_t at = d / 3;
    const int a = (int)(d - 3 * at);
    const double r0 = xyz[3 * at] - q0, r1 = xyz[3 * at + 1] - q1, r2 = xyz[3 * at + 2] - q2;
    const double w = a == 0 ? t0 + (f1 * r2 - f2 * r1) : (a == 1 ? t1 + (f2 * r0 - f0 * r2) : t2 + (f0 * r1 - f1 * r0));
    y[d] -= scale * w;
  }
}

// The update and the second projection of CGS2 in one read of V, per segment:
// u' = u - V[:J] h (h of the chunk's segment, row stride hld) and, per chunk c,
// part[c*(J+1) + j] = V[j, chunk] . u'[chunk] (j < J), part[c*(J+1) + J] = |u'[chunk]|^2.
// Tiles of C dofs of V[:J] are staged in LDS (J <= 4096 / C): 256 / C lanes per column form
// u' with a fixed butterfly, then thread j accumulates row j's dot over the tile.
template <int C>
__global__ void __launch_bounds__(256) k_seg_cgs_b(const int64_t *__restrict__ cbeg, const int64_t *__restrict__ cend,
                                                    const int32_t *__restrict__ cseg, int J,
                                                    const double *__restrict__ V, int64_t ld,
                                                    const double *__restrict__ h, int hld, double *__restrict__ u,
                                                    double *__restrict__ part) {
  constexpr int LPC = 256 / C;         // lanes per column
  constexpr int RS = 256 / C;          // rows staged per pass of the workgroup
  constexpr int NR = (4096 / C) / RS;  // staging loads per thread
  __shared__ double vs[4096 / C][C + 1];
  __shared__ double hs[4096 / C];
  __shared__ double up[C];
  const int tid = threadIdx.x;
  const int c = blockIdx.x;
  const int s = cseg[c];
  const int64_t b = cbeg[c], e = cend[c];
  for (int t = tid; t < J; t += 256) hs[t] = h[(int64_t)s * hld + t];
  double acc = 0.0, ss = 0.0;
  const int c_ld = tid % C, r_ld = tid / C;
  for (int64_t c0 = b; c0 < e; c0 += C) {
    __syncthreads();
    const int64_t il = c0 + c_ld;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const int r = r_ld + q * RS;
      if (r < J) vs[r][c_ld] = il < e ? V[(int64_t)r * ld + il] : 0.0;
    }
    __syncthreads();
    {
      const int col = tid / LPC, pl = tid % LPC;
      double sum = 0.0;
      for (int r = pl; r < J; r += LPC) sum += vs[r][col] * hs[r];
#pragma unroll
      for (int off = LPC / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, LPC);
      if (pl == 0) {
        const int64_t i = c0 + col;
        const double un = i < e ? u[i] - sum : 0.0;
        up[col] = un;
        if (i < e) u[i] = un;
      }
    }
    __syncthreads();
    if (tid < J) {
#pragma unroll
      for (int cc = 0; cc < C; ++cc) acc += vs[tid][cc] * up[cc];
    }
    if (tid == 0) {
#pragma unroll
      for (int cc = 0; cc < C; ++cc) ss += up[cc] * up[cc];
    }
  }
  if (tid < J) part[(int64_t)c * (J + 1) + tid] = acc;
  if (tid == 0) part[(int64_t)c * (J + 1) + J] = ss;
}

// Delayed CGS2 step, per segment (the segmented form of inch_dcgs2_step): i = J - 1, V[i] is
// the provisional vector, u = Op V[i].
// part[c*2J + j] = V_j . u and part[c*2J + J + j] = V_j . w over chunk c (w = V[J-1]);
// blockIdx.y picks the tile of 8 rows.
__global__ void __launch_bounds__(256) k_seg_dual_dot_part(const int64_t *__restrict__ cbeg,
                                                            const int64_t *__restrict__ cend, int J,
                                                            const double *__restrict__ V, int64_t ld,
                                                            const double *__restrict__ u,
                                                            const double *__restrict__ w, double *__restrict__ part) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int64_t b = cbeg[c], e = cend[c];
  {
    const int j0 = 8 * (int)blockIdx.y;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0}, g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
      const double ui = u[i], wi = w[i];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (j0 + t < J) {
          const double v = V[(int64_t)(j0 + t) * ld + i];
          a[t] += v * ui;
          g[t] += v * wi;
        }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const double sa = block_sum(a[t], red);
      const double sg = block_sum(g[t], red);
      if (threadIdx.x == 0 && j0 + t < J) {
        part[(int64_t)c * 2 * J + j0 + t] = sa;
        part[(int64_t)c * 2 * J + J + j0 + t] = sg;
      }
    }
  }
}

// One wave per segment s: sg[s] = (s_0..s_i, g_0..g_{i-1}, -). With v_i = ~v_i - sum g_l V_l:
// coef[s, i] = v_i . u = s_i - g . s, coef[s, j] = s_j - coef[s, i] g_j (j < i),
// coef[s, J + l] = g_l and alpha_i[s] = coef[s, i] - beta_prev[s] g_{i-1} (the one operator
// term left in a Krylov basis).
__global__ void __launch_bounds__(64) k_seg_dcgs2_coef(int J, const double *__restrict__ sg,
                                                        const double *__restrict__ beta_prev,
                                                        double *__restrict__ coef, double *__restrict__ alpha_i) {
  const int s = blockIdx.x, lane = threadIdx.x;
  const int i = J - 1;
  const double *sv = sg + (int64_t)s * 2 * J, *g = sv + J;
  double *cf = coef + (int64_t)s * 2 * J;
  double p = 0.0;
  for (int l = lane; l < i; l += 64) p += g[l] * sv[l];
  for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, 64);
  const double ci = sv[i] - p;
  for (int j = lane; j < i; j += 64) {
    cf[j] = sv[j] - ci * g[j];
    cf[J + j] = g[j];
  }
  if (lane == 0) {
    cf[i] = ci;
    alpha_i[s] = i > 0 ? ci - beta_prev[s] * g[i - 1] : ci;
  }
}

// u -= sum_{j<=i} coef[s, j] V_j and V[i] -= sum_{l<i} coef[s, J + l] V_l over chunk c of
// segment s, in one read of V[0..i]; part[c] = |u'|^2 over the chunk.
__global__ void __launch_bounds__(256) k_seg_dcgs2_update(const int64_t *__restrict__ cbeg,
                                                           const int64_t *__restrict__ cend,
                                                           const int32_t *__restrict__ cseg, int J,
                                                           double *__restrict__ V, int64_t ld,
                                                           const double *__restrict__ coef, double *__restrict__ u,
                                                           double *__restrict__ part) {
  __shared__ double ca[kSegMaxJ / 2], cg[kSegMaxJ / 2];
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int s = cseg[c];
  const int i = J - 1;
  const double *cf = coef + (int64_t)s * 2 * J;
  for (int t = threadIdx.x; t <= i; t += blockDim.x) ca[t] = cf[t];
  for (int t = threadIdx.x; t < i; t += blockDim.x) cg[t] = cf[J + t];
  __syncthreads();
  double ss = 0.0;
  for (int64_t x = cbeg[c] + threadIdx.x; x < cend[c]; x += blockDim.x) {
    double su = 0.0, sv = 0.0;
    for (int j = 0; j < i; ++j) {
      const double v = V[(int64_t)j * ld + x];
      su += ca[j] * v;
      sv += cg[j] * v;
    }
    const double vi = V[(int64_t)i * ld + x];
    const double un = u[x] - (su + ca[i] * vi);
    u[x] = un;
    V[(int64_t)i * ld + x] = vi - sv;
    ss += un * un;
  }
  const double tot = block_sum(ss, red);
  if (threadIdx.x == 0) part[c] = tot;
}

// dst[s] = sqrt(src[s * stride + off])
__global__ void k_seg_sqrt_col(int S, const double *__restrict__ src, int stride, int off, double *__restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < S) dst[t] = sqrt(src[(int64_t)t * stride + off]);
}

// dst[i] = src[i] / div[s] (dst2 too when given)
__global__ void __launch_bounds__(256) k_seg_scale(const int64_t *__restrict__ cbeg, const int64_t *__restrict__ cend,
                                                    const int32_t *__restrict__ cseg, const double *__restrict__ src,
                                                    const double *__restrict__ div, double *__restrict__ dst,
                                                    double *__restrict__ dst2) {
  const int c = blockIdx.x;
  const double inv = 1.0 / div[cseg[c]];
  for (int64_t i = cbeg[c] + threadIdx.x; i < cend[c]; i += blockDim.x) {
    const double v = src[i] * inv;
    dst[i] = v;
    if (dst2) dst2[i] = v;
  }
}

// out[q*ld_out + i] = sum_r W[s][r][q] * V[r*ld + i]  (W: [S, m, k] row-major), q < k <= KM
typedef double double4_t __attribute__((ext_vector_type(4)));

// out[q, i] = sum_r W[s][r][q] V[r, i] over chunk c (segment s), q < k <= 16 NT, with
// v_mfma_f64_16x16x4f64 (operand layouts as in k_rr_combine_mfma): W[s] is staged in LDS as
// [mp][16 NT + 1] (mp = m rounded up to 4, zero rows and columns beyond m and k); every wave
// takes 4 16-column slices of the chunk per pass (4 loads of V in flight) and keeps NT
// accumulator tiles of 16 x 16 for each.
template <int NT>
__global__ void __launch_bounds__(256) k_seg_combine_mfma(const int64_t *__restrict__ cbeg,
                                                           const int64_t *__restrict__ cend,
                                                           const int32_t *__restrict__ cseg, int m, int k,
                                                           const double *__restrict__ V, int64_t ld,
                                                           const double *__restrict__ W, double *__restrict__ out,
                                                           int64_t ld_out) {
  extern __shared__ double ws[];  // [mp][16 NT + 1]
  constexpr int KP = 16 * NT + 1;
  const int c = blockIdx.x;
  const int s = cseg[c];
  const int mp = (m + 3) & ~3;
  for (int t = threadIdx.x; t < mp * KP; t += blockDim.x) {
    const int r = t / KP, q = t % KP;
    ws[t] = (r < m && q < k) ? W[((int64_t)s * m + r) * k + q] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t e = cend[c];
  constexpr int NS = 4;  // 16-column slices per wave and pass (loads in flight)
  for (int64_t i0 = cbeg[c] + 16 * w; i0 < e; i0 += 64 * NS) {
    int64_t col[NS];
    double4_t acc[NS][NT];
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      col[s2] = i0 + 64 * s2 + (lane & 15);
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[s2][t] = (double4_t){0.0, 0.0, 0.0, 0.0};
    }
    for (int r0 = 0; r0 < mp; r0 += 4) {
      const int r = r0 + (lane >> 4);
      double bv[NS];
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) bv[s2] = (col[s2] < e && r < m) ? V[(int64_t)r * ld + col[s2]] : 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double av = ws[r * KP + 16 * t + (lane & 15)];
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2)
          acc[s2][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[s2], acc[s2][t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) {
      if (col[s2] >= e) continue;
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          const int q = 16 * t + (lane >> 4) + 4 * x;
          if (q < k) out[(int64_t)q * ld_out + col[s2]] = acc[s2][t][x];
        }
    }
  }
}

// part[c*k + q] = sum_{i in chunk} (Y[q,i] - theta[s*k+q] X[q,i])^2
__global__ void __launch_bounds__(256) k_seg_resnorm(const int64_t *__restrict__ cbeg, const int64_t *__restrict__ cend,
                                                      const int32_t *__restrict__ cseg, int k,
                                                      const double *__restrict__ X, int64_t ldx,
                                                      const double *__restrict__ Y, int64_t ldy,
                                                      const double *__restrict__ theta, double *__restrict__ part) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int s = cseg[c];
  for (int q = 0; q < k; ++q) {
    const double th = theta[(int64_t)s * k + q];
    double acc = 0.0;
    for (int64_t i = cbeg[c] + threadIdx.x; i < cend[c]; i += blockDim.x) {
      const double r = Y[(int64_t)q * ldy + i] - th * X[(int64_t)q * ldx + i];
      acc += r * r;
    }
    const double tot = block_sum(acc, red);
    if (threadIdx.x == 0) part[(int64_t)c * k + q] = tot;
  }
}

}  // namespace

extern "C" {

inch_status inch_seg_multidot(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend,
                              const int64_t *sptr, int32_t J, const double *V, int64_t ld, const double *u, double *h,
                              double *work, void *stream) {
  INCH_REQUIRE(nchunk > 0 && S > 0 && J > 0 && cbeg && cend && sptr && V && u && h && work,
               "inch_seg_multidot: bad args");
  hipStream_t st = (hipStream_t)stream;
  k_seg_dot_part<<<dim3(nchunk, (J + 7) / 8), 256, 0, st>>>(cbeg, cend, J, V, ld, u, work);
  HIP_TRY(hipGetLastError());
  const int64_t tot = (int64_t)S * J;
  k_seg_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(S, J, sptr, work, h, 0);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_seg_update(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg,
                            const int64_t *sptr, int32_t J, const double *V, int64_t ld, const double *h, int32_t hld,
                            double scale, double *u, double *nrm_out, double *work, void *stream) {
  INCH_REQUIRE(nchunk > 0 && S > 0 && J >= 0 && J <= kSegMaxJ && cbeg && cend && cseg && u,
               "inch_seg_update: bad args (J <= 512)");
  INCH_REQUIRE(J == 0 || (V && h && hld >= J), "inch_seg_update: V/h missing");
  INCH_REQUIRE(!nrm_out || (work && sptr), "inch_seg_update: norm needs work and sptr");
  hipStream_t st = (hipStream_t)stream;
  k_seg_update<<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, J, V, ld, h, hld, scale, u, nrm_out ? work : nullptr);
  HIP_TRY(hipGetLastError());
  if (nrm_out) {
    k_seg_reduce<<<(unsigned)((S + 255) / 256), 256, 0, st>>>(S, 1, sptr, work, nrm_out, 1);
    HIP_TRY(hipGetLastError());
  }
  return INCH_OK;
}

inch_status inch_seg_rigid_apply(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend,
                                 const int32_t *cseg, const int64_t *sptr, const double *xyz, const double *cen,
                                 const double *kinv, const double *ninv, const double *x, double *y, double scale,
                                 double *g, double *work, void *stream) {
  INCH_REQUIRE(nchunk > 0 && S > 0 && cbeg && cend && cseg && sptr && xyz && cen && x && g && work,
               "inch_seg_rigid_apply: bad args");
  INCH_REQUIRE(!y || (kinv && ninv), "inch_seg_rigid_apply: the update needs kinv and ninv");
  hipStream_t st = (hipStream_t)stream;
  k_seg_rigid_moments<<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, xyz, cen, x, work);
  HIP_TRY(hipGetLastError());
  k_seg_reduce<<<(unsigned)((6 * (int64_t)S + 255) / 256), 256, 0, st>>>(S, 6, sptr, work, g, 0);
  HIP_TRY(hipGetLastError());
  if (!y) return INCH_OK;
  k_seg_rigid_update<<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, xyz, cen, kinv, ninv, g, scale, y);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_seg_cgs_b(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg,
                           const int64_t *sptr, int32_t J, const double *V, int64_t ld, const double *h, int32_t hld,
                           double *u, double *h2, double *nrm_out, double *work, void *stream) {
  INCH_REQUIRE(nchunk > 0 && S > 0 && J > 0 && J <= 256 && hld >= J && cbeg && cend && cseg && sptr && V && h && u &&
                   h2 && nrm_out && work,
               "inch_seg_cgs_b: bad args (0 < J <= 256)");
  hipStream_t st = (hipStream_t)stream;
  if (J <= 64)
    k_seg_cgs_b<64><<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, J, V, ld, h, hld, u, work);
  else if (J <= 128)
    k_seg_cgs_b<32><<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, J, V, ld, h, hld, u, work);
  else
    k_seg_cgs_b<16><<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, J, V, ld, h, hld, u, work);
  HIP_TRY(hipGetLastError());
  const int64_t tot = (int64_t)S * (J + 1);
  k_seg_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(S, J + 1, sptr, work, h2, 0);
  HIP_TRY(hipGetLastError());
  k_seg_sqrt_col<<<(unsigned)((S + 255) / 256), 256, 0, st>>>(S, h2, J + 1, J, nrm_out);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

int64_t inch_seg_dcgs2_work_size(int32_t nchunk, int32_t S, int32_t J) {
  return (int64_t)nchunk * (2 * (int64_t)J + 1) + 4 * (int64_t)S * J;
}

inch_status inch_seg_dcgs2_step(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend,
                                const int32_t *cseg, const int64_t *sptr, int32_t J, double *V, int64_t ld, double *u,
                                double *alpha_i, const double *beta_prev, double *beta_out, double *work,
                                void *stream) {
  INCH_REQUIRE(nchunk > 0 && S > 0 && J > 0 && J <= kSegMaxJ / 2 && cbeg && cend && cseg && sptr && V && u &&
                   alpha_i && beta_out && work,
               "inch_seg_dcgs2_step: bad args (0 < J <= 256)");
  INCH_REQUIRE(J == 1 || beta_prev, "inch_seg_dcgs2_step: beta_prev needed for J > 1");
  hipStream_t st = (hipStream_t)stream;
  double *part = work;
  double *sg = part + (int64_t)nchunk * 2 * J;
  double *coef = sg + (int64_t)S * 2 * J;
  double *npart = coef + (int64_t)S * 2 * J;
  k_seg_dual_dot_part<<<dim3(nchunk, (J + 7) / 8), 256, 0, st>>>(cbeg, cend, J, V, ld, u, V + (int64_t)(J - 1) * ld,
                                                                 part);
  HIP_TRY(hipGetLastError());
  const int64_t tot = (int64_t)S * 2 * J;
  k_seg_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(S, 2 * J, sptr, part, sg, 0);
  HIP_TRY(hipGetLastError());
  k_seg_dcgs2_coef<<<S, 64, 0, st>>>(J, sg, beta_prev, coef, alpha_i);
  HIP_TRY(hipGetLastError());
  k_seg_dcgs2_update<<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, J, V, ld, coef, u, npart);
  HIP_TRY(hipGetLastError());
  k_seg_reduce<<<(unsigned)((S + 255) / 256), 256, 0, st>>>(S, 1, sptr, npart, beta_out, 1);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_seg_scale(int32_t nchunk, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg,
                           const double *src, const double *div, double *dst, double *dst2, void *stream) {
  INCH_REQUIRE(nchunk > 0 && cbeg && cend && cseg && src && div && dst, "inch_seg_scale: bad args");
  hipStream_t st = (hipStream_t)stream;
  k_seg_scale<<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, src, div, dst, dst2);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_seg_combine(int32_t nchunk, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg, int32_t m,
                             int32_t k, const double *V, int64_t ld, const double *W, double *out, int64_t ld_out,
                             void *stream) {
  INCH_REQUIRE(nchunk > 0 && cbeg && cend && cseg && V && W && out && m > 0 && k > 0 && k <= 64,
               "inch_seg_combine: bad args (k <= 64)");
  const int nt = k <= 16 ? 1 : (k <= 32 ? 2 : 4);
  const size_t lds = (size_t)((m + 3) & ~3) * (16 * nt + 1) * sizeof(double);
  INCH_REQUIRE(lds <= 64 * 1024, "inch_seg_combine: W[s] (m = %d, k = %d) does not fit the 64 KiB of LDS", m, k);
  hipStream_t st = (hipStream_t)stream;
  if (nt == 1)
    k_seg_combine_mfma<1><<<nchunk, 256, lds, st>>>(cbeg, cend, cseg, m, k, V, ld, W, out, ld_out);
  else if (nt == 2)
    k_seg_combine_mfma<2><<<nchunk, 256, lds, st>>>(cbeg, cend, cseg, m, k, V, ld, W, out, ld_out);
  else
    k_seg_combine_mfma<4><<<nchunk, 256, lds, st>>>(cbeg, cend, cseg, m, k, V, ld, W, out, ld_out);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

inch_status inch_seg_resnorm(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg,
                             const int64_t *sptr, int32_t k, const double *X, int64_t ldx, const double *Y,
                             int64_t ldy, const double *theta, double *out, double *work, void *stream) {
  INCH_REQUIRE(nchunk > 0 && S > 0 && k > 0 && cbeg && cend && cseg && sptr && X && Y && theta && out && work,
               "inch_seg_resnorm: bad args");
  hipStream_t st = (hipStream_t)stream;
  k_seg_resnorm<<<nchunk, 256, 0, st>>>(cbeg, cend, cseg, k, X, ldx, Y, ldy, theta, work);
  HIP_TRY(hipGetLastError());
  const int64_t tot = (int64_t)S * k;
  k_seg_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(S, k, sptr, work, out, 1);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

}  // extern "C"
