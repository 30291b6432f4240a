// Exclusive prefix sums on the device, shared by the construction and the ordering:
// out[i] = sum in[0..i-1] for i <= m (out[m] = the total). One block of 1024 threads walks
// short arrays in chunks; longer ones take three passes: per-tile sums (kScanTile elements
// per 256-thread block), the single-block scan of the tile sums, and a rescan of every tile
// from its offset.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

template <typename T>
__global__ void k_scan_one(int64_t m, const T *__restrict__ in, T *__restrict__ out) {
  __shared__ T buf[1024];
  __shared__ T carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < m; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const T v = i < m ? in[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const T t = (int)threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < m) out[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = carry;
}

constexpr int kScanPer = 16;                 // elements per thread of a tile
constexpr int kScanTile = 256 * kScanPer;    // elements per tile (one 256-thread block)

template <typename T>
__global__ void __launch_bounds__(256) k_scan_tile_sums(int64_t m, const T *__restrict__ in, T *__restrict__ sums) {
  __shared__ T red[256];
  const int64_t b0 = blockIdx.x * (int64_t)kScanTile;
  T s = 0;
  for (int q = threadIdx.x; q < kScanTile; q += 256) {
    const int64_t i = b0 + q;
    if (i < m) s += in[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

template <typename T>
__global__ void __launch_bounds__(256) k_scan_tiles(int64_t m, const T *__restrict__ in, const T *__restrict__ offs,
                                                    T *__restrict__ out) {
  __shared__ T ts[256];
  const int64_t i0 = blockIdx.x * (int64_t)kScanTile + threadIdx.x * (int64_t)kScanPer;
  T v[kScanPer];
  T tot = 0;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    v[q] = i0 + q < m ? in[i0 + q] : 0;
    tot += v[q];
  }
  ts[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const T t = (int)threadIdx.x >= off ? ts[threadIdx.x - off] : 0;
    __syncthreads();
    ts[threadIdx.x] += t;
    __syncthreads();
  }
  T run = offs[blockIdx.x] + ts[threadIdx.x] - tot;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    if (i0 + q < m) out[i0 + q] = run;
    run += v[q];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[m] = offs[gridDim.x];
}

inline int64_t scan_tiles(int64_t m) { return (m + kScanTile - 1) / kScanTile; }

// The scan of in[0..m) into out[0..m] on stream st, no allocation: sums and offs hold
// scan_tiles(m) + 1 elements each (unused, and may be null, when scan_tiles(m) <= 4).
template <typename T>
hipError_t scan_exclusive_into(int64_t m, const T *in, T *out, T *sums, T *offs, hipStream_t st) {
  const int64_t nt = scan_tiles(m);
  if (nt <= 4) {
    k_scan_one<T><<<1, 1024, 0, st>>>(m, in, out);
    return hipGetLastError();
  }
  k_scan_tile_sums<T><<<(unsigned)nt, 256, 0, st>>>(m, in, sums);
  k_scan_one<T><<<1, 1024, 0, st>>>(nt, sums, offs);
  k_scan_tiles<T><<<(unsigned)nt, 256, 0, st>>>(m, in, offs, out);
  return hipGetLastError();
}

}  // namespace
