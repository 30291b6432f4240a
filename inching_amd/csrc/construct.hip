// Construction kernels for gfx950: cell-list neighbour search and ANM 3x3
// super-element assembly into full-storage block rows.
//
// Neighbour rule (shared with assembly and the matrix-free operator):
//   j in ball(i)  <=>  (dx*dx + dy*dy) + dz*dz <= rc*rc,  dx = x_j - x_i (float64),
// evaluated with explicit round-to-nearest operations so no FMA contraction can
// change the set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "scan.h"

#include <hipcub/hipcub.hpp>

namespace {

constexpr int kMaxRow = 2048;  // capacity of one neighbour row in the sort kernel

__device__ __forceinline__ double dist2(double dx, double dy, double dz) {
  return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

struct Grid {
  double lo[3];
  double inv;  // 1 / cell size
  int64_t dim[3];
};

__device__ __forceinline__ int64_t cell_coord(double v, double lo, double inv, int64_t dim) {
  int64_t c = (int64_t)floor((v - lo) * inv);
  return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

// Minimum and maximum that propagate NaN (fmin/fmax drop it), so that a NaN coordinate
// reaches the host's finiteness check.
__device__ __forceinline__ double nan_min(double a, double b) { return (a < b || a != a) ? a : b; }
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

__global__ void k_minmax(int64_t n, const double *__restrict__ xyz, double *__restrict__ part) {
  // part[blockIdx.x*6 + {0..2}] = min, {3..5} = max; NaN in either when any coordinate is NaN
  __shared__ double smin[3][256], smax[3][256];
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    for (int a = 0; a < 3; ++a) {
      double v = xyz[3 * i + a];
      mn[a] = nan_min(mn[a], v);
      mx[a] = nan_max(mx[a], v);
    }
  }
  for (int a = 0; a < 3; ++a) {
    smin[a][threadIdx.x] = mn[a];
    smax[a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int a = 0; a < 3; ++a) {
        smin[a][threadIdx.x] = nan_min(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
        smax[a][threadIdx.x] = nan_max(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 3; ++a) {
      part[blockIdx.x * 6 + a] = smin[a][0];
      part[blockIdx.x * 6 + 3 + a] = smax[a][0];
    }
}

__global__ void k_cell_count(int64_t n, const double *__restrict__ xyz, Grid g, int32_t *__restrict__ cell_of,
                             int32_t *__restrict__ count) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t c[3];
  for (int a = 0; a < 3; ++a) c[a] = cell_coord(xyz[3 * i + a], g.lo[a], g.inv, g.dim[a]);
  int64_t id = (c[0] * g.dim[1] + c[1]) * g.dim[2] + c[2];
  cell_of[i] = (int32_t)id;
  atomicAdd(&count[id], 1);
}


__global__ void k_cell_fill(int64_t n, const int32_t *__restrict__ cell_of, const int32_t *__restrict__ start,
                            int32_t *__restrict__ cursor, int32_t *__restrict__ atoms) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c = cell_of[i];
  int32_t p = atomicAdd(&cursor[c], 1);
  atoms[start[c] + p] = (int32_t)i;
}

// One thread per atom: count (indices == nullptr) or write neighbours in cell order.
__global__ void k_neigh(int64_t n, const double *__restrict__ xyz, Grid g, double rc2,
                        const int32_t *__restrict__ start, const int32_t *__restrict__ atoms,
                        const int64_t *__restrict__ indptr, int64_t *__restrict__ counts,
                        int32_t *__restrict__ indices) {
  // threads take the atoms in cell order, so that the lanes of a wave scan the same cells
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t i = atoms[t];
  const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  int64_t c0 = cell_coord(xi, g.lo[0], g.inv, g.dim[0]);
  int64_t c1 = cell_coord(yi, g.lo[1], g.inv, g.dim[1]);
  int64_t c2 = cell_coord(zi, g.lo[2], g.inv, g.dim[2]);
  int64_t cnt = 0;
  int64_t out = indices ? indptr[i] : 0;
  int64_t end = indices ? indptr[i + 1] : 0;
  for (int64_t a = c0 - 1; a <= c0 + 1; ++a) {
    if (a < 0 || a >= g.dim[0]) continue;
    for (int64_t b = c1 - 1; b <= c1 + 1; ++b) {
      if (b < 0 || b >= g.dim[1]) continue;
      for (int64_t c = c2 - 1; c <= c2 + 1; ++c) {
        if (c < 0 || c >= g.dim[2]) continue;
        int64_t id = (a * g.dim[1] + b) * g.dim[2] + c;
        for (int32_t p = start[id]; p < start[id + 1]; ++p) {
          int32_t j = atoms[p];
          double d2 = dist2(__dsub_rn(xyz[3 * (int64_t)j], xi), __dsub_rn(xyz[3 * (int64_t)j + 1], yi),
                            __dsub_rn(xyz[3 * (int64_t)j + 2], zi));
          if (d2 <= rc2) {
            if (indices) {
              if (out < end) indices[out] = j;
              ++out;
            }
            ++cnt;
          }
        }
      }
    }
  }
  if (!indices) counts[i] = cnt;
}

// Periodic images in x and y (membrane boxes): image s = (sx, sy) in {-1,0,1}^2 shifts atom j
// by (sx * bx, sy * by, 0); bit (sx + 1) * 3 + (sy + 1) of an entry's mask marks an image
// within the cutoff, d_s = (x_j + s box) - x_i. Cells are periodic in x and y (cell size
// >= rc, coordinates wrapped into the box from its origin) and open in z.
struct PbcGrid {
  double lo[3];      // box origin in x, y; minimum z
  double cell[3];    // cell sizes
  double box[2];
  int64_t dim[3];
};

__device__ __forceinline__ int64_t pcell(double v, double lo, double box, double cell, int64_t dim) {
  double w = v - lo;
  w -= box * floor(w / box);
  int64_t c = (int64_t)floor(w / cell);
  return c < 0 ? 0 : (c >= dim ? dim - 1 : c);
}

__device__ __forceinline__ int64_t zcell(double v, const PbcGrid &g) {
  int64_t c = (int64_t)floor((v - g.lo[2]) / g.cell[2]);
  return c < 0 ? 0 : (c >= g.dim[2] ? g.dim[2] - 1 : c);
}

__device__ __forceinline__ uint32_t image_mask(const double *__restrict__ xyz, int64_t i, int64_t j, double bx,
                                               double by, double rc2) {
  const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  const double dz = __dsub_rn(xyz[3 * j + 2], zi);
  uint32_t m = 0;
  for (int sx = -1; sx <= 1; ++sx) {
    const double dx = __dsub_rn(__dadd_rn(xyz[3 * j], sx * bx), xi);
    for (int sy = -1; sy <= 1; ++sy) {
      const double dy = __dsub_rn(__dadd_rn(xyz[3 * j + 1], sy * by), yi);
      if (dist2(dx, dy, dz) <= rc2) m |= 1u << ((sx + 1) * 3 + (sy + 1));
    }
  }
  return m;
}

__global__ void k_cell_count_pbc(int64_t n, const double *__restrict__ xyz, PbcGrid g, int32_t *__restrict__ cell_of,
                                 int32_t *__restrict__ count) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t c0 = pcell(xyz[3 * i], g.lo[0], g.box[0], g.cell[0], g.dim[0]);
  const int64_t c1 = pcell(xyz[3 * i + 1], g.lo[1], g.box[1], g.cell[1], g.dim[1]);
  const int64_t c2 = zcell(xyz[3 * i + 2], g);
  const int64_t id = (c0 * g.dim[1] + c1) * g.dim[2] + c2;
  cell_of[i] = (int32_t)id;
  atomicAdd(&count[id], 1);
}

// One thread per atom; every cell is visited once even when a periodic dimension has fewer
// than three cells; an entry is kept when any image is within the cutoff.
__global__ void k_neigh_pbc(int64_t n, const double *__restrict__ xyz, PbcGrid g, double rc2,
                            const int32_t *__restrict__ start, const int32_t *__restrict__ atoms,
                            const int64_t *__restrict__ indptr, int64_t *__restrict__ counts,
                            int32_t *__restrict__ indices) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;   // atoms in cell order
  if (t >= n) return;
  const int64_t i = atoms[t];
  const int64_t c0 = pcell(xyz[3 * i], g.lo[0], g.box[0], g.cell[0], g.dim[0]);
  const int64_t c1 = pcell(xyz[3 * i + 1], g.lo[1], g.box[1], g.cell[1], g.dim[1]);
  const int64_t c2 = zcell(xyz[3 * i + 2], g);
  const int64_t na = g.dim[0] < 3 ? g.dim[0] : 3, nbb = g.dim[1] < 3 ? g.dim[1] : 3;
  int64_t cnt = 0;
  int64_t out = indices ? indptr[i] : 0;
  const int64_t end = indices ? indptr[i + 1] : 0;
  for (int64_t ia = 0; ia < na; ++ia) {
    const int64_t a = g.dim[0] < 3 ? ia : (c0 - 1 + ia + g.dim[0]) % g.dim[0];
    for (int64_t ib = 0; ib < nbb; ++ib) {
      const int64_t b = g.dim[1] < 3 ? ib : (c1 - 1 + ib + g.dim[1]) % g.dim[1];
      for (int64_t c = c2 - 1; c <= c2 + 1; ++c) {
        if (c < 0 || c >= g.dim[2]) continue;
        const int64_t id = (a * g.dim[1] + b) * g.dim[2] + c;
        for (int32_t p = start[id]; p < start[id + 1]; ++p) {
          const int32_t j = atoms[p];
          if (image_mask(xyz, i, j, g.box[0], g.box[1], rc2)) {
            if (indices) {
              if (out < end) indices[out] = j;
              ++out;
            }
            ++cnt;
          }
        }
      }
    }
  }
  if (!indices) counts[i] = cnt;
}

__global__ void k_image_masks(int64_t n, const double *__restrict__ xyz, double bx, double by, double rc2,
                              const int64_t *__restrict__ indptr, const int32_t *__restrict__ indices,
                              uint16_t *__restrict__ images) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) images[e] = (uint16_t)image_mask(xyz, i, indices[e], bx, by, rc2);
}

// Sorted copy of a row held in LDS (buf[0..len)) into out[0..len), ascending, by one block
// (the caller has synchronised after filling buf). Rows of up to 256 entries by rank (each
// entry's place = the number of smaller entries, ties by position); longer ones by a
// bitonic sort in place over the next power of two (buf holds kMaxRow entries).
__device__ void sort_row_to(int32_t *buf, int len, int32_t *__restrict__ out) {
  const int tid = threadIdx.x, nt = blockDim.x;
  if (len <= 256) {
    for (int t = tid; t < len; t += nt) {
      const int32_t v = buf[t];
      int r = 0;
      for (int q = 0; q < len; ++q) {
        const int32_t w = buf[q];
        r += (w < v) | ((w == v) & (q < t));
      }
      out[r] = v;
    }
    return;
  }
  int p2 = 1;
  while (p2 < len) p2 <<= 1;
  for (int t = len + tid; t < p2; t += nt) buf[t] = 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= p2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = tid; t < p2; t += nt) {
        const int ixj = t ^ j;
        if (ixj > t) {
          const bool up = (t & k) == 0;
          const int32_t x = buf[t], y = buf[ixj];
          if ((x > y) == up) {
            buf[t] = y;
            buf[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int t = tid; t < len; t += nt) out[t] = buf[t];
}

// One 64-thread block per row: the row's columns sorted ascending in place.
__global__ void __launch_bounds__(64) k_sort_rows(int64_t n, const int64_t *__restrict__ indptr,
                                                   int32_t *__restrict__ indices, int32_t *__restrict__ overflow) {
  __shared__ int32_t buf[kMaxRow];
  const int64_t row = blockIdx.x;
  if (row >= n) return;
  const int64_t b = indptr[row], e = indptr[row + 1];
  const int len = (int)(e - b);
  if (len > kMaxRow) {
    if (threadIdx.x == 0) atomicMax(overflow, len);
    return;
  }
  for (int t = threadIdx.x; t < len; t += blockDim.x) buf[t] = indices[b + t];
  __syncthreads();
  sort_row_to(buf, len, indices + b);
}

// counts[i] = length of row order[i] of the source CSR
__global__ void k_perm_counts(int64_t n, const int64_t *__restrict__ indptr, const int32_t *__restrict__ order,
                              int64_t *__restrict__ counts) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = order[i];
  counts[i] = indptr[o + 1] - indptr[o];
}

// One block per new row: the columns of source row order[row], renumbered by undo (old index
// -> new index), sorted ascending (sort_row_to).
__global__ void __launch_bounds__(64) k_permute_sort_rows(int64_t n, const int64_t *__restrict__ indptr,
                                                           const int32_t *__restrict__ indices,
                                                           const int32_t *__restrict__ order,
                                                           const int64_t *__restrict__ undo,
                                                           const int64_t *__restrict__ out_ptr,
                                                           int32_t *__restrict__ out_idx,
                                                           int32_t *__restrict__ overflow) {
  __shared__ int32_t buf[kMaxRow];
  const int64_t row = blockIdx.x;
  if (row >= n) return;
  const int64_t o = order[row];
  const int64_t b = indptr[o];
  const int len = (int)(indptr[o + 1] - b);
  if (len > kMaxRow) {
    if (threadIdx.x == 0) atomicMax(overflow, len);
    return;
  }
  for (int t = threadIdx.x; t < len; t += blockDim.x) buf[t] = (int32_t)undo[indices[b + t]];
  __syncthreads();
  sort_row_to(buf, len, out_idx + out_ptr[row]);
}

// One 64-lane wave per row: H_ij = -d d^T / r^2 for j != i (r^2 > 0); the diagonal
// block is minus the sum of the row's off-diagonal blocks plus plus_i * I.
__global__ void __launch_bounds__(256) k_assemble(int64_t n, const double *__restrict__ xyz, double plus_i,
                                                   const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                   const uint16_t *__restrict__ images, double bx, double by,
                                                   double *__restrict__ v64, float *__restrict__ v32,
                                                   double *__restrict__ diag) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  const double xi = xyz[3 * row], yi = xyz[3 * row + 1], zi = xyz[3 * row + 2];
  double s[6] = {0, 0, 0, 0, 0, 0};  // xx xy xz yy yz zz of sum d d^T / r^2
  int64_t selfpos = -1;
  const int64_t b0 = rowptr[row], b1 = rowptr[row + 1];
  for (int64_t b = b0 + lane; b < b1; b += 64) {
    int32_t j = col[b];
    double h[9];
    if (j == row) {
      selfpos = b;
      continue;
    }
    if (!images) {
      double dx = __dsub_rn(xyz[3 * (int64_t)j], xi), dy = __dsub_rn(xyz[3 * (int64_t)j + 1], yi),
             dz = __dsub_rn(xyz[3 * (int64_t)j + 2], zi);
      double r2 = dist2(dx, dy, dz);
      double c = r2 > 0.0 ? -1.0 / r2 : 0.0;
      double d[3] = {dx, dy, dz};
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) h[3 * p + q] = d[p] * d[q] * c;
    } else {
      // sum over the images of j within the cutoff
      for (int t = 0; t < 9; ++t) h[t] = 0.0;
      const uint32_t m = images[b];
      const double dz = __dsub_rn(xyz[3 * (int64_t)j + 2], zi);
      for (int bit = 0; bit < 9; ++bit) {
        if (!((m >> bit) & 1u)) continue;
        const double sx = (double)(bit / 3 - 1), sy = (double)(bit % 3 - 1);
        const double dx = __dsub_rn(__dadd_rn(xyz[3 * (int64_t)j], sx * bx), xi);
        const double dy = __dsub_rn(__dadd_rn(xyz[3 * (int64_t)j + 1], sy * by), yi);
        const double r2 = dist2(dx, dy, dz);
        const double c = r2 > 0.0 ? -1.0 / r2 : 0.0;
        const double d[3] = {dx, dy, dz};
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 3; ++q) h[3 * p + q] += d[p] * d[q] * c;
      }
    }
    s[0] -= h[0];
    s[1] -= h[1];
    s[2] -= h[2];
    s[3] -= h[4];
    s[4] -= h[5];
    s[5] -= h[8];
    if (v64)
      for (int t = 0; t < 9; ++t) v64[9 * b + t] = h[t];
    if (v32)
      for (int t = 0; t < 9; ++t) v32[9 * b + t] = (float)h[t];
  }
  // deterministic butterfly reduction over the wave
  for (int off = 32; off > 0; off >>= 1)
    for (int t = 0; t < 6; ++t) s[t] += __shfl_xor(s[t], off, 64);
  // the lane holding the self block writes it
  if (selfpos >= 0) {
    double h[9] = {s[0] + plus_i, s[1], s[2], s[1], s[3] + plus_i, s[4], s[2], s[4], s[5] + plus_i};
    if (v64)
      for (int t = 0; t < 9; ++t) v64[9 * selfpos + t] = h[t];
    if (v32)
      for (int t = 0; t < 9; ++t) v32[9 * selfpos + t] = (float)h[t];
  }
  if (diag && lane == 0) {
    diag[3 * row] = s[0] + plus_i;
    diag[3 * row + 1] = s[3] + plus_i;
    diag[3 * row + 2] = s[5] + plus_i;
  }
}

// ---- cluster layout of the matrix-free operator --------------------------------------
// Atoms are grouped into clusters of kClusterAtoms that are compact in space: cells of
// cell_xy x cell_xy x cell_z, numbered column-major (x, y column, then z inside the column),
// atoms in a cell by index, cut into groups of 4; the groups are numbered by their smallest
// neighbour index, so consecutive clusters gather from one band of the vectors. Each cluster
// lists the union of its atoms' neighbour rows once, as (j, mask) with mask bit t set when j
// is a neighbour of the cluster's atom t (j itself excluded), so the product gathers a
// neighbour's coordinates and vector entries once per cluster instead of once per atom.
constexpr int kClusterAtoms = 4;
constexpr int kMaxUnion = 4096;  // row entries one cluster may gather (LDS sort capacity)

struct CGrid {
  double lo[3];
  double inv[3];  // 1 / cell size per axis
  int64_t dim[3];
};

__global__ void k_cl_cell(int64_t n, const double *__restrict__ xyz, CGrid g, int32_t *__restrict__ cell_of,
                          int32_t *__restrict__ count) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t c[3];
  for (int a = 0; a < 3; ++a) c[a] = cell_coord(xyz[3 * i + a], g.lo[a], g.inv[a], g.dim[a]);
  const int64_t id = (c[0] * g.dim[1] + c[1]) * g.dim[2] + c[2];
  cell_of[i] = (int32_t)id;
  atomicAdd(&count[id], 1);
}

// One thread per cell: insertion sort of the cell's atoms by index (cells hold a few atoms),
// which makes the order independent of the atomic fill order.
__global__ void k_cell_isort(int64_t ncell, const int32_t *__restrict__ start, int32_t *__restrict__ atoms) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int32_t b = start[c], e = start[c + 1];
  for (int32_t p = b + 1; p < e; ++p) {
    const int32_t v = atoms[p];
    int32_t q = p - 1;
    while (q >= b && atoms[q] > v) {
      atoms[q + 1] = atoms[q];
      --q;
    }
    atoms[q + 1] = v;
  }
}

__global__ void k_cl_atoms(int64_t n, int64_t slots, const int32_t *__restrict__ atoms,
                           int32_t *__restrict__ cluster_atoms) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < slots) cluster_atoms[q] = q < n ? atoms[q] : -1;
}

// Squared distance of atoms i and j as the cluster product computes it (d = p_j - p_i, the
// same FMAs): a pair at r2 = 0 (coincident atoms, whose term the reference masks) gets no
// list bit, so the product needs no test of its own.
__device__ __forceinline__ double cl_pair_r2(const double *__restrict__ xyz, int64_t i, int64_t j) {
  const double d0 = xyz[3 * j] - xyz[3 * i], d1 = xyz[3 * j + 1] - xyz[3 * i + 1], d2 = xyz[3 * j + 2] - xyz[3 * i + 2];
  return __builtin_fma(d2, d2, __builtin_fma(d1, d1, d0 * d0));
}

// One block per cluster. The rows of its atoms (each in ascending column order, as the
// neighbour search and the renumbering write them) are merged into keys (j << 3 | t) in
// ascending order in LDS, and each run of equal j becomes one entry with the OR of its atom
// bits.
// Count pass (FILL = false): counts[c] = entries, base[c] = smallest j; overflow bit 1 =
// more than kMaxUnion row entries, bit 2 = the union spans 2^24 columns or more. With
// prov_ptr given it also writes the entries, at prov_ptr[c] (room for all of the cluster's row
// entries), for a later copy into the final layout.
// Fill pass: entries[cl_ptr[c] + r] = ((j - base[c]) << 8) | mask, ascending j.
template <bool FILL>
__global__ void __launch_bounds__(256) k_cl_union(int64_t n, int64_t ncl, const double *__restrict__ xyz,
                                                   const int64_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col,
                                                   const int32_t *__restrict__ cluster_atoms,
                                                   const int64_t *__restrict__ cl_ptr, int32_t *__restrict__ base,
                                                   int64_t *__restrict__ counts, uint32_t *__restrict__ entries,
                                                   int32_t *__restrict__ overflow,
                                                   const int64_t *__restrict__ prov_ptr = nullptr) {
  __shared__ uint64_t key[kMaxUnion];
  __shared__ int32_t val[kMaxUnion];
  __shared__ int32_t part[256];
  __shared__ int64_t rb[kClusterAtoms], rl[kClusterAtoms + 1];
  __shared__ int32_t ca[kClusterAtoms];
  __shared__ int32_t s_min, s_max, s_bad;
  const int64_t c = blockIdx.x;
  const int tid = threadIdx.x;
  if (tid == 0) {
    int64_t acc = 0;
    for (int t = 0; t < kClusterAtoms; ++t) {
      const int32_t a = cluster_atoms[kClusterAtoms * c + t];
      ca[t] = a;
      const int64_t b = a >= 0 ? rowptr[a] : 0, e = a >= 0 ? rowptr[a + 1] : 0;
      rb[t] = b;
      rl[t] = acc;
      acc += e - b;
    }
    rl[kClusterAtoms] = acc;
    s_min = 0;
    s_max = 0;
    s_bad = 0;
  }
  __syncthreads();
  const int64_t total64 = rl[kClusterAtoms];
  if (total64 > kMaxUnion) {  // uniform over the block
    if (tid == 0) {
      atomicOr(overflow, 1);
      if (!FILL) counts[c] = 0;
    }
    return;
  }
  const int total = (int)total64;
  // the four rows (ascending columns) into val[]; each entry's place in the merged order is
  // its index in its own row plus, for every other row, the entries below it (equal columns
  // of earlier rows count as below), found by binary search: the keys (j << 3 | t) come out
  // in ascending order, as a sort of them would give. A row out of order sets overflow bit 4.
  for (int e = tid; e < total; e += blockDim.x) {
    int t = 0;
    while (e >= rl[t + 1]) ++t;
    val[e] = col[rb[t] + (e - rl[t])];
  }
  __syncthreads();
  // rows out of order would make the ranks collide and leave keys unwritten, and a column
  // outside [0, n) would address coordinates out of bounds: the block stops here, before any
  // rank is formed (uniform exit)
  for (int e = tid; e < total; e += blockDim.x) {
    int t = 0;
    while (e >= rl[t + 1]) ++t;
    if ((e > rl[t] && val[e - 1] >= val[e]) || val[e] < 0 || val[e] >= n) s_bad = 1;
  }
  __syncthreads();
  if (s_bad) {
    if (tid == 0) {
      atomicOr(overflow, 4);
      if (!FILL) counts[c] = 0;
    }
    return;
  }
  for (int e = tid; e < total; e += blockDim.x) {
    int t = 0;
    while (e >= rl[t + 1]) ++t;
    const int32_t j = val[e];
    int rank = e - (int)rl[t];
    for (int u = 0; u < kClusterAtoms; ++u) {
      if (u == t) continue;
      int lo = (int)rl[u], hi = (int)rl[u + 1];
      while (lo < hi) {
        const int m = (lo + hi) >> 1;
        const int32_t x = val[m];
        if (u < t ? x <= j : x < j)
          lo = m + 1;
        else
          hi = m;
      }
      rank += lo - (int)rl[u];
    }
    key[rank] = ((uint64_t)(uint32_t)j << 3) | (uint64_t)t;
  }
  __syncthreads();
  // runs of equal j: the run's first element carries the mask (0 = nothing to emit)
  const int chunk = (total + 255) / 256;
  const int e0 = tid * chunk, e1 = min(total, e0 + chunk);
  int cnt = 0;
  for (int e = e0; e < e1; ++e) {
    const uint32_t j = (uint32_t)(key[e] >> 3);
    int32_t m = 0;
    if (e == 0 || (uint32_t)(key[e - 1] >> 3) != j) {
      for (int f = e; f < total && (uint32_t)(key[f] >> 3) == j; ++f) {
        const int t = (int)(key[f] & 7);
        if ((uint32_t)ca[t] != j && cl_pair_r2(xyz, ca[t], j) > 0.0) m |= 1 << t;
      }
    }
    val[e] = m;
    cnt += m != 0;
  }
  part[tid] = cnt;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  const int before = part[tid] - cnt, emitted = part[255];
  if (!FILL) {
    if (cnt > 0) {
      int first = -1, last = -1;
      for (int e = e0; e < e1; ++e)
        if (val[e]) {
          if (first < 0) first = e;
          last = e;
        }
      if (before == 0) s_min = (int32_t)(key[first] >> 3);
      if (before + cnt == emitted) s_max = (int32_t)(key[last] >> 3);
    }
    __syncthreads();
    if (tid == 0) {
      counts[c] = emitted;
      base[c] = s_min;
      if (emitted > 0 && (int64_t)s_max - (int64_t)s_min >= ((int64_t)1 << 24)) atomicOr(overflow, 2);
    }
    if (prov_ptr) {
      const int64_t out = prov_ptr[c] + before;
      const uint32_t b0 = (uint32_t)s_min;
      int r = 0;
      for (int e = e0; e < e1; ++e)
        if (val[e]) entries[out + r++] = (((uint32_t)(key[e] >> 3) - b0) << 8) | (uint32_t)val[e];
    }
    return;
  }
  const int64_t out = cl_ptr[c] + before;
  const uint32_t b0 = (uint32_t)base[c];
  int r = 0;
  for (int e = e0; e < e1; ++e)
    if (val[e]) entries[out + r++] = (((uint32_t)(key[e] >> 3) - b0) << 8) | (uint32_t)val[e];
}

// Sort keys of the clusters: smallest neighbour in the high word, cluster index in the low.
__global__ void k_cl_keys(int64_t ncl, const int32_t *__restrict__ base, uint64_t *__restrict__ key) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c < ncl) key[c] = ((uint64_t)(uint32_t)base[c] << 32) | (uint64_t)c;
}

// Clusters renumbered: slot c' takes cluster perm[c'] = low word of the c'-th sorted key
// (atoms, list length, base).
__global__ void k_cl_permute(int64_t ncl, const uint64_t *__restrict__ perm, const int32_t *__restrict__ atoms_in,
                             const int64_t *__restrict__ counts_in, const int32_t *__restrict__ base_in,
                             int32_t *__restrict__ atoms_out, int64_t *__restrict__ counts_out,
                             int32_t *__restrict__ base_out) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  const int64_t src = (int64_t)(perm[c] & 0xffffffffu);
  for (int t = 0; t < kClusterAtoms; ++t) atoms_out[kClusterAtoms * c + t] = atoms_in[kClusterAtoms * src + t];
  counts_out[c] = counts_in[src];
  base_out[c] = base_in[src];
}

// tot[c] = entries of the rows of cluster c's atoms (room for its list before the union)
__global__ void k_cl_totals(int64_t ncl, const int32_t *__restrict__ cluster_atoms,
                            const int64_t *__restrict__ rowptr, int64_t *__restrict__ tot) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  int64_t s = 0;
  for (int t = 0; t < kClusterAtoms; ++t) {
    const int32_t a = cluster_atoms[kClusterAtoms * c + t];
    if (a >= 0) s += rowptr[a + 1] - rowptr[a];
  }
  tot[c] = s;
}

// Final layout from the count pass's provisional lists: cluster c' (new numbering) takes the
// list of cluster perm[c'] (old numbering), written at prov_ptr of the old cluster.
__global__ void __launch_bounds__(256) k_cl_copy(int64_t ncl, const uint64_t *__restrict__ perm,
                                                 const int64_t *__restrict__ prov_ptr,
                                                 const uint32_t *__restrict__ prov,
                                                 const int64_t *__restrict__ cl_ptr, uint32_t *__restrict__ entries) {
  const int64_t c = blockIdx.x;
  const int64_t src = prov_ptr[(int64_t)(perm[c] & 0xffffffffu)];
  const int64_t dst = cl_ptr[c], len = cl_ptr[c + 1] - dst;
  for (int64_t e = threadIdx.x; e < len; e += blockDim.x) entries[dst + e] = prov[src + e];
}

// ---- record form of the cluster layout (the operator's Chebyshev recurrence) ----
// rec[6 q .. 6 q + 2] = coordinates of slot q = 4 c + t (atom cluster_atoms[q]; zeros for a
// padding slot), rec[6 q + 3 .. 6 q + 5] = 0 (the vector part, written by the pack and the
// product); slot_of[atom] = its slot.
__global__ void k_rec_coords(int64_t slots, const int32_t *__restrict__ cluster_atoms, const double *__restrict__ xyz,
                             double *__restrict__ rec, int32_t *__restrict__ slot_of) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= slots) return;
  const int32_t at = cluster_atoms[q];
  for (int k = 0; k < 3; ++k) {
    rec[6 * q + k] = at >= 0 ? xyz[3 * (int64_t)at + k] : 0.0;
    rec[6 * q + 3 + k] = 0.0;
  }
  if (at >= 0) slot_of[at] = (int32_t)q;
}

// One block per cluster: its list rewritten as 48 (slot of j) | mask -- the byte offset of
// j's 48-byte record, whose low 4 bits are zero, carrying the mask -- and sorted ascending
// (bitonic sort in LDS; a list holds fewer than kMaxUnion entries).
__global__ void __launch_bounds__(256) k_rec_lists(int64_t ncl, const int64_t *__restrict__ cl_ptr,
                                                    const int32_t *__restrict__ base,
                                                    const uint32_t *__restrict__ entries,
                                                    const int32_t *__restrict__ slot_of, uint32_t *__restrict__ out) {
  __shared__ uint32_t key[kMaxUnion];
  const int64_t c = blockIdx.x;
  const int64_t e0 = cl_ptr[c];
  const int len = (int)(cl_ptr[c + 1] - e0);
  const uint32_t b0 = (uint32_t)base[c];
  int p2 = 1;
  while (p2 < len) p2 <<= 1;
  for (int e = threadIdx.x; e < p2; e += blockDim.x) {
    uint32_t k = 0xffffffffu;
    if (e < len) {
      const uint32_t w = entries[e0 + e];
      k = 48u * (uint32_t)slot_of[b0 + (w >> 8)] | (w & 15u);
    }
    key[e] = k;
  }
  __syncthreads();
  for (int k = 2; k <= p2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < p2; t += blockDim.x) {
        const int ixj = t ^ j;
        if (ixj > t) {
          const bool up = (t & k) == 0;
          const uint32_t a = key[t], b = key[ixj];
          if ((a > b) == up) {
            key[t] = b;
            key[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < len; e += blockDim.x) out[e0 + e] = key[e];
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) INCH_FAIL(INCH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

namespace {

// out[0..m] = exclusive scan of in[0..m-1] (out[m] = total), on stream st; the tile
// temporaries are allocated here and freed after the scan has finished.
template <typename T>
inch_status scan_exclusive(int64_t m, const T *in, T *out, hipStream_t st) {
  const int64_t nt = scan_tiles(m);
  DevBuf sums, offs;
  if (nt > 4) {
    HIP_TRY(hipMalloc(&sums.p, (nt + 1) * sizeof(T)));
    HIP_TRY(hipMalloc(&offs.p, (nt + 1) * sizeof(T)));
  }
  HIP_TRY(scan_exclusive_into(m, in, out, (T *)sums.p, (T *)offs.p, st));
  if (nt > 4) HIP_TRY(hipStreamSynchronize(st));
  return INCH_OK;
}

}  // namespace

#define SCAN_TRY(m, in, out)                                   \
  do {                                                         \
    const inch_status _s = scan_exclusive((m), (in), (out), st); \
    if (_s != INCH_OK) return _s;                              \
  } while (0)

extern "C" inch_status inch_neighbors_celllist(int64_t n, const double *xyz, double rc, int64_t *indptr,
                                               int32_t *indices, int64_t capacity, int64_t *nnz_out,
                                               void *stream) {
  INCH_REQUIRE(n > 0 && xyz && indptr && rc > 0, "inch_neighbors_celllist: bad args");
  INCH_REQUIRE(n < (int64_t)1 << 31, "inch_neighbors_celllist: n must fit int32");
  hipStream_t st = (hipStream_t)stream;
  // bounds
  const int nb = 256;
  DevBuf part;
  HIP_TRY(hipMalloc(&part.p, nb * 6 * sizeof(double)));
  k_minmax<<<nb, 256, 0, st>>>(n, xyz, (double *)part.p);
  HIP_TRY(hipGetLastError());
  std::vector<double> hp(nb * 6);
  HIP_TRY(hipMemcpyAsync(hp.data(), part.p, nb * 6 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  Grid g;
  double hi[3];
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = INFINITY;
    hi[a] = -INFINITY;
    bool nan = false;   // std::min/max would drop a NaN partial
    for (int b = 0; b < nb; ++b) {
      nan |= std::isnan(hp[b * 6 + a]) || std::isnan(hp[b * 6 + 3 + a]);
      g.lo[a] = std::min(g.lo[a], hp[b * 6 + a]);
      hi[a] = std::max(hi[a], hp[b * 6 + 3 + a]);
    }
    if (nan || !std::isfinite(g.lo[a]) || !std::isfinite(hi[a])) INCH_FAIL(INCH_ERR_NAN, "inch_neighbors_celllist: non-finite coordinates");
  }
  g.inv = 1.0 / rc;
  int64_t ncell = 1;
  for (int a = 0; a < 3; ++a) {
    g.dim[a] = std::max<int64_t>(1, (int64_t)std::floor((hi[a] - g.lo[a]) * g.inv) + 1);
    ncell *= g.dim[a];
  }
  if (ncell > ((int64_t)1 << 30)) INCH_FAIL(INCH_ERR_CAPACITY, "inch_neighbors_celllist: %lld cells", (long long)ncell);
  DevBuf cell_of, cnt, start, atoms, counts, ovf;
  HIP_TRY(hipMalloc(&cell_of.p, n * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&cnt.p, (ncell + 1) * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&start.p, (ncell + 1) * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&atoms.p, n * sizeof(int32_t)));
  HIP_TRY(hipMemsetAsync(cnt.p, 0, (ncell + 1) * sizeof(int32_t), st));
  const int bs = 256;
  const int64_t gb = (n + bs - 1) / bs;
  k_cell_count<<<gb, bs, 0, st>>>(n, xyz, g, (int32_t *)cell_of.p, (int32_t *)cnt.p);
  HIP_TRY(hipGetLastError());
  SCAN_TRY(ncell, (const int32_t *)cnt.p, (int32_t *)start.p);
  HIP_TRY(hipMemsetAsync(cnt.p, 0, (ncell + 1) * sizeof(int32_t), st));
  k_cell_fill<<<gb, bs, 0, st>>>(n, (int32_t *)cell_of.p, (int32_t *)start.p, (int32_t *)cnt.p, (int32_t *)atoms.p);
  HIP_TRY(hipGetLastError());
  const double rc2 = rc * rc;
  if (indices == nullptr) {
    HIP_TRY(hipMalloc(&counts.p, n * sizeof(int64_t)));
    k_neigh<<<gb, bs, 0, st>>>(n, xyz, g, rc2, (int32_t *)start.p, (int32_t *)atoms.p, nullptr, (int64_t *)counts.p,
                               nullptr);
    HIP_TRY(hipGetLastError());
    SCAN_TRY(n, (const int64_t *)counts.p, indptr);
    int64_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (nnz_out) *nnz_out = total;
    return INCH_OK;
  }
  int64_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (total > capacity) INCH_FAIL(INCH_ERR_CAPACITY, "inch_neighbors_celllist: need %lld > capacity %lld", (long long)total, (long long)capacity);
  k_neigh<<<gb, bs, 0, st>>>(n, xyz, g, rc2, (int32_t *)start.p, (int32_t *)atoms.p, indptr, nullptr, indices);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMalloc(&ovf.p, sizeof(int32_t)));
  HIP_TRY(hipMemsetAsync(ovf.p, 0, sizeof(int32_t), st));
  k_sort_rows<<<n, 64, 0, st>>>(n, indptr, indices, (int32_t *)ovf.p);
  HIP_TRY(hipGetLastError());
  int32_t over = 0;
  HIP_TRY(hipMemcpyAsync(&over, ovf.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (over) INCH_FAIL(INCH_ERR_CAPACITY, "inch_neighbors_celllist: a row has %d neighbours (> %d)", over, kMaxRow);
  if (nnz_out) *nnz_out = total;
  return INCH_OK;
}

extern "C" inch_status inch_csr_permute(int64_t n, const int64_t *indptr, const int32_t *indices,
                                        const int32_t *order, const int64_t *undo, int64_t *out_indptr,
                                        int32_t *out_indices, void *stream) {
  INCH_REQUIRE(n > 0 && indptr && indices && order && undo && out_indptr && out_indices, "inch_csr_permute: bad args");
  hipStream_t st = (hipStream_t)stream;
  DevBuf counts, ovf;
  HIP_TRY(hipMalloc(&counts.p, n * sizeof(int64_t)));
  HIP_TRY(hipMalloc(&ovf.p, sizeof(int32_t)));
  HIP_TRY(hipMemsetAsync(ovf.p, 0, sizeof(int32_t), st));
  const int bs = 256;
  k_perm_counts<<<(unsigned)((n + bs - 1) / bs), bs, 0, st>>>(n, indptr, order, (int64_t *)counts.p);
  HIP_TRY(hipGetLastError());
  SCAN_TRY(n, (const int64_t *)counts.p, out_indptr);
  k_permute_sort_rows<<<(unsigned)n, 64, 0, st>>>(n, indptr, indices, order, undo, out_indptr, out_indices,
                                                  (int32_t *)ovf.p);
  HIP_TRY(hipGetLastError());
  int32_t over = 0;
  HIP_TRY(hipMemcpyAsync(&over, ovf.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (over) INCH_FAIL(INCH_ERR_CAPACITY, "inch_csr_permute: a row has %d entries (> %d)", over, kMaxRow);
  return INCH_OK;
}

extern "C" inch_status inch_hessian_assemble(int64_t n, const double *xyz, double rc, double plus_i,
                                             const int64_t *rowptr, const int32_t *col, double *vals64,
                                             float *vals32, double *diag, void *stream) {
  (void)rc;  // the pattern already encodes the cutoff; only r^2 == 0 is masked here
  INCH_REQUIRE(n > 0 && xyz && rowptr && col, "inch_hessian_assemble: bad args");
  INCH_REQUIRE(vals64 || vals32 || diag, "inch_hessian_assemble: no output");
  hipStream_t st = (hipStream_t)stream;
  const int bs = 256;
  const int64_t grid = (n * 64 + bs - 1) / bs;
  k_assemble<<<grid, bs, 0, st>>>(n, xyz, plus_i, rowptr, col, nullptr, 0.0, 0.0, vals64, vals32, diag);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

extern "C" inch_status inch_neighbors_celllist_pbc(int64_t n, const double *xyz, double rc, double box_x,
                                                   double box_y, double origin_x, double origin_y, int64_t *indptr,
                                                   int32_t *indices, uint16_t *images, int64_t capacity,
                                                   int64_t *nnz_out, void *stream) {
  INCH_REQUIRE(n > 0 && xyz && indptr && rc > 0, "inch_neighbors_celllist_pbc: bad args");
  INCH_REQUIRE(box_x > rc && box_y > rc, "ABORTED. The PBC box size must exceed the cutoff (%g, %g vs %g)", box_x,
               box_y, rc);
  INCH_REQUIRE(n < (int64_t)1 << 31, "inch_neighbors_celllist_pbc: n must fit int32");
  INCH_REQUIRE(!indices || images, "inch_neighbors_celllist_pbc: the fill pass needs images");
  hipStream_t st = (hipStream_t)stream;
  const int nb = 256;
  DevBuf part;
  HIP_TRY(hipMalloc(&part.p, nb * 6 * sizeof(double)));
  k_minmax<<<nb, 256, 0, st>>>(n, xyz, (double *)part.p);
  HIP_TRY(hipGetLastError());
  std::vector<double> hp(nb * 6);
  HIP_TRY(hipMemcpyAsync(hp.data(), part.p, nb * 6 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  double zlo = INFINITY, zhi = -INFINITY;
  for (int b = 0; b < nb; ++b) {
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(hp[b * 6 + a]) && hp[b * 6 + a] != INFINITY)
        INCH_FAIL(INCH_ERR_NAN, "inch_neighbors_celllist_pbc: non-finite coordinates");
    zlo = std::min(zlo, hp[b * 6 + 2]);
    zhi = std::max(zhi, hp[b * 6 + 5]);
  }
  if (!std::isfinite(zlo) || !std::isfinite(zhi)) INCH_FAIL(INCH_ERR_NAN, "inch_neighbors_celllist_pbc: non-finite coordinates");
  PbcGrid g;
  g.lo[0] = origin_x;
  g.lo[1] = origin_y;
  g.lo[2] = zlo;
  g.box[0] = box_x;
  g.box[1] = box_y;
  g.dim[0] = std::max<int64_t>(1, (int64_t)std::floor(box_x / rc));
  g.dim[1] = std::max<int64_t>(1, (int64_t)std::floor(box_y / rc));
  g.cell[0] = box_x / (double)g.dim[0];
  g.cell[1] = box_y / (double)g.dim[1];
  g.cell[2] = rc;
  g.dim[2] = std::max<int64_t>(1, (int64_t)std::floor((zhi - zlo) / rc) + 1);
  const int64_t ncell = g.dim[0] * g.dim[1] * g.dim[2];
  if (ncell > ((int64_t)1 << 30)) INCH_FAIL(INCH_ERR_CAPACITY, "inch_neighbors_celllist_pbc: %lld cells", (long long)ncell);
  DevBuf cell_of, cnt, start, atoms, counts, ovf;
  HIP_TRY(hipMalloc(&cell_of.p, n * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&cnt.p, (ncell + 1) * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&start.p, (ncell + 1) * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&atoms.p, n * sizeof(int32_t)));
  HIP_TRY(hipMemsetAsync(cnt.p, 0, (ncell + 1) * sizeof(int32_t), st));
  const int bs = 256;
  const int64_t gb = (n + bs - 1) / bs;
  k_cell_count_pbc<<<gb, bs, 0, st>>>(n, xyz, g, (int32_t *)cell_of.p, (int32_t *)cnt.p);
  HIP_TRY(hipGetLastError());
  SCAN_TRY(ncell, (const int32_t *)cnt.p, (int32_t *)start.p);
  HIP_TRY(hipMemsetAsync(cnt.p, 0, (ncell + 1) * sizeof(int32_t), st));
  k_cell_fill<<<gb, bs, 0, st>>>(n, (int32_t *)cell_of.p, (int32_t *)start.p, (int32_t *)cnt.p, (int32_t *)atoms.p);
  HIP_TRY(hipGetLastError());
  const double rc2 = rc * rc;
  if (indices == nullptr) {
    HIP_TRY(hipMalloc(&counts.p, n * sizeof(int64_t)));
    k_neigh_pbc<<<gb, bs, 0, st>>>(n, xyz, g, rc2, (int32_t *)start.p, (int32_t *)atoms.p, nullptr,
                                   (int64_t *)counts.p, nullptr);
    HIP_TRY(hipGetLastError());
    SCAN_TRY(n, (const int64_t *)counts.p, indptr);
    int64_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (nnz_out) *nnz_out = total;
    return INCH_OK;
  }
  int64_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (total > capacity) INCH_FAIL(INCH_ERR_CAPACITY, "inch_neighbors_celllist_pbc: need %lld > capacity %lld", (long long)total, (long long)capacity);
  k_neigh_pbc<<<gb, bs, 0, st>>>(n, xyz, g, rc2, (int32_t *)start.p, (int32_t *)atoms.p, indptr, nullptr, indices);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMalloc(&ovf.p, sizeof(int32_t)));
  HIP_TRY(hipMemsetAsync(ovf.p, 0, sizeof(int32_t), st));
  k_sort_rows<<<n, 64, 0, st>>>(n, indptr, indices, (int32_t *)ovf.p);
  HIP_TRY(hipGetLastError());
  int32_t over = 0;
  HIP_TRY(hipMemcpyAsync(&over, ovf.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (over) INCH_FAIL(INCH_ERR_CAPACITY, "inch_neighbors_celllist_pbc: a row has %d neighbours (> %d)", over, kMaxRow);
  k_image_masks<<<gb, bs, 0, st>>>(n, xyz, box_x, box_y, rc2, indptr, indices, images);
  HIP_TRY(hipGetLastError());
  if (nnz_out) *nnz_out = total;
  return INCH_OK;
}

extern "C" inch_status inch_hessian_assemble_pbc(int64_t n, const double *xyz, double plus_i, const int64_t *rowptr,
                                                 const int32_t *col, const uint16_t *images, double box_x,
                                                 double box_y, double *vals64, float *vals32, double *diag,
                                                 void *stream) {
  INCH_REQUIRE(n > 0 && xyz && rowptr && col && images && box_x > 0 && box_y > 0, "inch_hessian_assemble_pbc: bad args");
  INCH_REQUIRE(vals64 || vals32 || diag, "inch_hessian_assemble_pbc: no output");
  hipStream_t st = (hipStream_t)stream;
  const int bs = 256;
  const int64_t grid = (n * 64 + bs - 1) / bs;
  k_assemble<<<grid, bs, 0, st>>>(n, xyz, plus_i, rowptr, col, images, box_x, box_y, vals64, vals32, diag);
  HIP_TRY(hipGetLastError());
  return INCH_OK;
}

// The cluster layout, in two calls (one_pass = false: the first, without entries, builds the
// clusters and their counts; the second, with entries, merges the rows again in the final
// numbering) or in one (the count pass also writes each list into a provisional buffer with
// room for the cluster's row entries, and the lists are copied into the final numbering).
static inch_status anm_clusters_impl(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                                     double cell_xy, double cell_z, int32_t *cluster_atoms, int64_t *cl_ptr,
                                     int32_t *base, uint32_t *entries, int64_t capacity, int64_t *nnz_out,
                                     void *stream, bool one_pass) {
  INCH_REQUIRE(n > 0 && xyz && rowptr && col && cluster_atoms && cl_ptr && base && cell_xy > 0 && cell_z > 0,
               "inch_anm_clusters: bad args");
  INCH_REQUIRE(n < (int64_t)1 << 31, "inch_anm_clusters: n must fit int32");
  // the cluster kernels address coordinates and vectors with 32-bit byte offsets
  if (24 * n > (int64_t)UINT32_MAX) INCH_FAIL(INCH_ERR_CAPACITY, "inch_anm_clusters: n = %lld exceeds the 32-bit offset range (%lld atoms)", (long long)n, (long long)(UINT32_MAX / 24));
  hipStream_t st = (hipStream_t)stream;
  const int64_t ncl = (n + kClusterAtoms - 1) / kClusterAtoms;
  DevBuf ovf;
  HIP_TRY(hipMalloc(&ovf.p, sizeof(int32_t)));
  HIP_TRY(hipMemsetAsync(ovf.p, 0, sizeof(int32_t), st));
  auto check_overflow = [&](const char *pass) -> inch_status {
    int32_t over = 0;
    HIP_TRY(hipMemcpyAsync(&over, ovf.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (over & 1) INCH_FAIL(INCH_ERR_CAPACITY, "inch_anm_clusters (%s): a cluster's rows hold more than %d entries", pass, kMaxUnion);
    if (over & 2) INCH_FAIL(INCH_ERR_CAPACITY, "inch_anm_clusters (%s): a cluster's neighbours span 2^24 columns or more", pass);
    if (over & 4) INCH_FAIL(INCH_ERR_INVALID, "inch_anm_clusters (%s): a row's columns are not in ascending order or outside [0, n)", pass);
    return INCH_OK;
  };
  if (entries && !one_pass) {  // fill pass on the layout of the count pass
    int64_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, cl_ptr + ncl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (total > capacity) INCH_FAIL(INCH_ERR_CAPACITY, "inch_anm_clusters: need %lld > capacity %lld", (long long)total, (long long)capacity);
    k_cl_union<true><<<(unsigned)ncl, 256, 0, st>>>(n, ncl, xyz, rowptr, col, cluster_atoms, cl_ptr, base, nullptr, entries,
                                                    (int32_t *)ovf.p);
    HIP_TRY(hipGetLastError());
    inch_status s = check_overflow("fill");
    if (s != INCH_OK) return s;
    if (nnz_out) *nnz_out = total;
    return INCH_OK;
  }
  // bounds
  const int nb = 256;
  DevBuf part;
  HIP_TRY(hipMalloc(&part.p, nb * 6 * sizeof(double)));
  k_minmax<<<nb, 256, 0, st>>>(n, xyz, (double *)part.p);
  HIP_TRY(hipGetLastError());
  std::vector<double> hp(nb * 6);
  HIP_TRY(hipMemcpyAsync(hp.data(), part.p, nb * 6 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  CGrid g;
  double hi[3];
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = INFINITY;
    hi[a] = -INFINITY;
    bool nan = false;   // std::min/max would drop a NaN partial
    for (int b = 0; b < nb; ++b) {
      nan |= std::isnan(hp[b * 6 + a]) || std::isnan(hp[b * 6 + 3 + a]);
      g.lo[a] = std::min(g.lo[a], hp[b * 6 + a]);
      hi[a] = std::max(hi[a], hp[b * 6 + 3 + a]);
    }
    if (nan || !std::isfinite(g.lo[a]) || !std::isfinite(hi[a])) INCH_FAIL(INCH_ERR_NAN, "inch_anm_clusters: non-finite coordinates");
  }
  // cells: cell_xy in x and y; cell_z in z, coarsened only beyond max(16n, 65536) cells (an
  // assembly's bounding box is mostly empty: coarser cells would make its clusters taller)
  const double size[3] = {cell_xy, cell_xy, cell_z};
  for (int a = 0; a < 3; ++a) {
    g.inv[a] = 1.0 / size[a];
    g.dim[a] = std::max<int64_t>(1, (int64_t)std::floor((hi[a] - g.lo[a]) * g.inv[a]) + 1);
  }
  const int64_t cap_cells = std::max<int64_t>(16 * n, 65536);
  while (g.dim[0] * g.dim[1] * g.dim[2] > cap_cells && g.dim[2] > 1) {
    g.inv[2] *= 0.5;
    g.dim[2] = std::max<int64_t>(1, (int64_t)std::floor((hi[2] - g.lo[2]) * g.inv[2]) + 1);
  }
  const int64_t ncell = g.dim[0] * g.dim[1] * g.dim[2];
  if (ncell > ((int64_t)1 << 30)) INCH_FAIL(INCH_ERR_CAPACITY, "inch_anm_clusters: %lld cells", (long long)ncell);
  DevBuf cell_of, cnt, start, atoms, counts;
  HIP_TRY(hipMalloc(&cell_of.p, n * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&cnt.p, (ncell + 1) * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&start.p, (ncell + 1) * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&atoms.p, n * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&counts.p, ncl * sizeof(int64_t)));
  HIP_TRY(hipMemsetAsync(cnt.p, 0, (ncell + 1) * sizeof(int32_t), st));
  const int bs = 256;
  const int64_t gb = (n + bs - 1) / bs;
  k_cl_cell<<<gb, bs, 0, st>>>(n, xyz, g, (int32_t *)cell_of.p, (int32_t *)cnt.p);
  HIP_TRY(hipGetLastError());
  SCAN_TRY(ncell, (const int32_t *)cnt.p, (int32_t *)start.p);
  HIP_TRY(hipMemsetAsync(cnt.p, 0, (ncell + 1) * sizeof(int32_t), st));
  k_cell_fill<<<gb, bs, 0, st>>>(n, (int32_t *)cell_of.p, (int32_t *)start.p, (int32_t *)cnt.p, (int32_t *)atoms.p);
  HIP_TRY(hipGetLastError());
  k_cell_isort<<<(unsigned)((ncell + bs - 1) / bs), bs, 0, st>>>(ncell, (int32_t *)start.p, (int32_t *)atoms.p);
  HIP_TRY(hipGetLastError());
  const int64_t slots = ncl * kClusterAtoms;
  k_cl_atoms<<<(unsigned)((slots + bs - 1) / bs), bs, 0, st>>>(n, slots, (int32_t *)atoms.p, cluster_atoms);
  HIP_TRY(hipGetLastError());
  DevBuf atoms0, counts0, base0, perm;
  HIP_TRY(hipMalloc(&atoms0.p, slots * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&counts0.p, ncl * sizeof(int64_t)));
  HIP_TRY(hipMalloc(&base0.p, ncl * sizeof(int32_t)));
  HIP_TRY(hipMalloc(&perm.p, ncl * sizeof(uint64_t)));
  HIP_TRY(hipMemcpyAsync(atoms0.p, cluster_atoms, slots * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  DevBuf prov_ptr, prov;
  if (one_pass) {   // room for every cluster's row entries, in the count pass's numbering
    HIP_TRY(hipMalloc(&prov_ptr.p, (ncl + 1) * sizeof(int64_t)));
    k_cl_totals<<<(unsigned)((ncl + bs - 1) / bs), bs, 0, st>>>(ncl, (int32_t *)atoms0.p, rowptr,
                                                               (int64_t *)counts0.p);
    HIP_TRY(hipGetLastError());
    SCAN_TRY(ncl, (const int64_t *)counts0.p, (int64_t *)prov_ptr.p);
    int64_t room = 0;
    HIP_TRY(hipMemcpyAsync(&room, (int64_t *)prov_ptr.p + ncl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMalloc(&prov.p, std::max<int64_t>(1, room) * sizeof(uint32_t)));
  }
  k_cl_union<false><<<(unsigned)ncl, 256, 0, st>>>(n, ncl, xyz, rowptr, col, (int32_t *)atoms0.p, nullptr, (int32_t *)base0.p,
                                                   (int64_t *)counts0.p, (uint32_t *)prov.p, (int32_t *)ovf.p,
                                                   (const int64_t *)prov_ptr.p);
  HIP_TRY(hipGetLastError());
  {
    inch_status s0 = check_overflow("count");
    if (s0 != INCH_OK) return s0;
  }
  // clusters in the order of their smallest neighbour (ties by spatial order): consecutive
  // clusters, and so each XCD's contiguous share of them, gather from one band of the
  // vectors, which are in the matrix's (RCM) order. Radix sort of the keys on the device.
  DevBuf keys, tmp;
  HIP_TRY(hipMalloc(&keys.p, ncl * sizeof(uint64_t)));
  k_cl_keys<<<(unsigned)((ncl + bs - 1) / bs), bs, 0, st>>>(ncl, (int32_t *)base0.p, (uint64_t *)keys.p);
  HIP_TRY(hipGetLastError());
  size_t tmp_bytes = 0;
  HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, (const uint64_t *)keys.p, (uint64_t *)perm.p,
                                            (int)ncl, 0, 64, st));
  HIP_TRY(hipMalloc(&tmp.p, tmp_bytes));
  HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, (const uint64_t *)keys.p, (uint64_t *)perm.p, (int)ncl,
                                            0, 64, st));
  k_cl_permute<<<(unsigned)((ncl + bs - 1) / bs), bs, 0, st>>>(ncl, (const uint64_t *)perm.p, (int32_t *)atoms0.p,
                                                              (int64_t *)counts0.p, (int32_t *)base0.p, cluster_atoms,
                                                              (int64_t *)counts.p, base);
  HIP_TRY(hipGetLastError());
  SCAN_TRY(ncl, (const int64_t *)counts.p, cl_ptr);
  inch_status s = check_overflow("count");
  if (s != INCH_OK) return s;
  int64_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, cl_ptr + ncl, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (nnz_out) *nnz_out = total;
  if (one_pass) {
    if (total > capacity) INCH_FAIL(INCH_ERR_CAPACITY, "inch_anm_clusters_one_pass: need %lld > capacity %lld", (long long)total, (long long)capacity);
    k_cl_copy<<<(unsigned)ncl, 256, 0, st>>>(ncl, (const uint64_t *)perm.p, (const int64_t *)prov_ptr.p,
                                             (const uint32_t *)prov.p, cl_ptr, entries);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));   // prov is freed on return
  }
  return INCH_OK;
}

extern "C" inch_status inch_anm_clusters(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                                         double cell_xy, double cell_z, int32_t *cluster_atoms, int64_t *cl_ptr,
                                         int32_t *base, uint32_t *entries, int64_t capacity, int64_t *nnz_out,
                                         void *stream) {
  return anm_clusters_impl(n, xyz, rowptr, col, cell_xy, cell_z, cluster_atoms, cl_ptr, base, entries, capacity,
                           nnz_out, stream, false);
}

extern "C" inch_status inch_anm_clusters_one_pass(int64_t n, const double *xyz, const int64_t *rowptr,
                                                  const int32_t *col, double cell_xy, double cell_z,
                                                  int32_t *cluster_atoms, int64_t *cl_ptr, int32_t *base,
                                                  uint32_t *entries, int64_t capacity, int64_t *nnz_out,
                                                  void *stream) {
  INCH_REQUIRE(entries && capacity >= 0, "inch_anm_clusters_one_pass: entries needed");
  return anm_clusters_impl(n, xyz, rowptr, col, cell_xy, cell_z, cluster_atoms, cl_ptr, base, entries, capacity,
                           nnz_out, stream, true);
}

extern "C" inch_status inch_anm_records(int64_t n, const double *xyz, int64_t ncl, const int32_t *cluster_atoms,
                                        const int64_t *cl_ptr, const int32_t *base, const uint32_t *entries,
                                        double *rec, uint32_t *rec_entries, void *stream) {
  INCH_REQUIRE(n > 0 && xyz && cluster_atoms && cl_ptr && base && entries && rec && rec_entries,
               "inch_anm_records: bad args");
  INCH_REQUIRE(ncl == (n + kClusterAtoms - 1) / kClusterAtoms, "inch_anm_records: ncl must be ceil(n / 4)");
  // the record product addresses records with 32-bit byte offsets (48 bytes per slot)
  if (192 * ncl > (int64_t)UINT32_MAX)
    INCH_FAIL(INCH_ERR_CAPACITY, "inch_anm_records: %lld clusters exceed the 32-bit record offsets", (long long)ncl);
  hipStream_t st = (hipStream_t)stream;
  const int64_t slots = kClusterAtoms * ncl;
  DevBuf slot_of;
  HIP_TRY(hipMalloc(&slot_of.p, n * sizeof(int32_t)));
  k_rec_coords<<<(unsigned)((slots + 255) / 256), 256, 0, st>>>(slots, cluster_atoms, xyz, rec, (int32_t *)slot_of.p);
  HIP_TRY(hipGetLastError());
  k_rec_lists<<<(unsigned)ncl, 256, 0, st>>>(ncl, cl_ptr, base, entries, (const int32_t *)slot_of.p, rec_entries);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));   // slot_of is freed on return
  return INCH_OK;
}
