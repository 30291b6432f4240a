/*
 * inching.h -- C ABI of libinching_hip.so, the MI355X (gfx950) implementation of
 * Inching's eigenproblem construction + diagonalisation path.
 *
 * Conventions
 *  - Every entry point returns inch_status (0 = OK, negative = error).
 *  - Device pointers are plain pointers to memory the caller owns (allocated by
 *    the caller, e.g. through torch); the library never frees caller memory.
 *  - `stream` is a hipStream_t passed as void*; NULL means the default stream.
 *    Launches are asynchronous on that stream unless the comment says "syncs".
 *  - Atoms are in the caller's units (the reference uses nm with the cutoff in
 *    Angstrom / 10); dof n = 3 * n_atoms, vectors interleave x,y,z per atom.
 *  - Block matrices ("block rows") are full-storage 3x3 block CSR:
 *      rowptr int64[n_atoms+1], col int32[nnzb] (ascending per row, diagonal
 *      included), vals T[nnzb*9] row-major 3x3 per block.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *  inch_neighbors_celllist   scipy cKDTree.query_ball_point calls of
 *                            InchingLiteInteger/Fuel/Coordinate/T1.py:219,239,276-277,365,397
 *  inch_rcm_host, inch_rcm_device
 *                            X_KdCuthillMckeeOrder, Fuel/Coordinate/T1.py:171-323
 *  inch_batch_head           GetPartitionTree/FlattenPartitionTree util.py:563-587 and
 *                            Fuel/T1.py:145-153
 *  inch_hessian_assemble     Xnumpy_SparseCupyMatrixUngappped.ReturnCupyHLowerTriangleInt32/
 *                            Int64, Fuel/T1.py:228-660 / :666-1021
 *  inch_bsr_spmv             OOC2_HalfMemS_v_KrylovAv_VOID / OOC2_FullMemS_v_KrylovAv_VOID
 *                            closures, Burn/Krylov/T3.py:54-239 (plus the Hotelling term of
 *                            Burn/Orthogonalization/T3.py:62-164 and the Chebyshev step of
 *                            Burn/PolynomialFilters/T2.py:59-80 fused as epilogues)
 *  inch_multidot / inch_gemv_update / inch_vec_*
 *                            cuBLAS dot/nrm2/gemv calls of
 *                            Burn/ThickRestartLanczosHotellingDeflation/T1.py:465-520 and
 *                            Burn/Orthogonalization/T2.py:64-137
 *  inch_rr_combine           V[:k] = W^T V, TRLM/T1.py:176,246; IRLM/T1.py:286-294; JDM/T1.py:553
 *  inch_anm_matfree_apply    the residual check of the drivers,
 *                            Command8A/commandBenchmark_InchingTRLMHD_Linux_Pdb_0064_A100.py:470-476
 */
#ifndef INCHING_H
#define INCHING_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t inch_status;
enum {
  INCH_OK = 0,
  INCH_ERR_INVALID = -1,      /* bad argument / shape */
  INCH_ERR_HIP = -2,          /* a HIP runtime call failed */
  INCH_ERR_CAPACITY = -3,     /* an output buffer or a fixed capacity is too small */
  INCH_ERR_NOT_CONVERGED = -4,
  INCH_ERR_NAN = -5
};

/* dtype codes */
enum { INCH_F32 = 0, INCH_F64 = 1 };

const char *inch_status_string(inch_status s);
const char *inch_last_error(void);   /* last error message of this thread */
int32_t inch_abi_version(void);

/* ---------------------------------------------------------------- host ---- */

/* Seeded random-sequential-addition globule: n atoms uniform in a sphere of the
 * given radius centred at the origin with pairwise distance >= dmin (cell grid,
 * xoshiro256** stream seeded by splitmix64(seed)). Writes xyz[n*3] (float64).
 * Returns INCH_ERR_CAPACITY if the sphere saturates before n atoms are placed
 * (after max_attempts darts). Host only. */
inch_status inch_synth_globule(int64_t n, double radius, double dmin, uint64_t seed,
                               int64_t max_attempts, double *xyz);

/* Reverse Cuthill-McKee on a neighbour CSR (self included, ascending columns).
 * seed_order: optional int64[n] seed order (NULL = stable argsort of degree).
 * order[n] (int32) = new->old; undo[n] (int64) = argsort(order). Host only. */
inch_status inch_rcm_host(int64_t n, const int64_t *indptr, const int32_t *indices,
                          const int64_t *seed_order, int32_t reverse, int32_t *order,
                          int64_t *undo);

/* The same permutation computed on the device (level-synchronous BFS, bit-exact to
 * inch_rcm_host; replaces X_KdCuthillMckeeOrder, Fuel/Coordinate/T1.py:171-323).
 * indptr / indices / seeds / order / undo are device pointers; rows hold unique column
 * indices (the neighbour CSR). seeds: int64[n] seed order, a permutation (the stable
 * argsort of the degree reproduces inch_rcm_host's default). work: device scratch of
 * inch_rcm_device_work_size(n) bytes. Syncs the stream (one level size per BFS level
 * comes back to the host). Inputs are validated on the device before use. Connected
 * components are contiguous in the order; when comp_ptr (host, comp_cap entries) is given it
 * receives their start offsets followed by n, and *n_comp (host, optional) their number. */
int64_t inch_rcm_device_work_size(int64_t n);
inch_status inch_rcm_device(int64_t n, const int64_t *indptr, const int32_t *indices, const int64_t *seeds,
                            int32_t reverse, int32_t *order, int64_t *undo, void *work, int64_t work_bytes,
                            int64_t *comp_ptr, int64_t comp_cap, int64_t *n_comp, void *stream);

/* Row-batch heads of recursive halving with leaves <= maxleaf, sorted descending
 * and cumulated: head[0..nbatch]. cap = capacity of head. Host only. */
inch_status inch_batch_head(int64_t n, int64_t maxleaf, int64_t *head, int64_t cap,
                            int64_t *nbatch);

/* -------------------------------------------------------------- device ---- */

/* Cell-list neighbour search on the device. xyz: device float64[n*3].
 * Phase 1 (indices == NULL): writes indptr[n+1] (int64, device) and returns the
 *   total count in *nnz_out (syncs).
 * Phase 2 (indices != NULL, capacity >= nnz): fills indices (int32, ascending
 *   per row, self included). Neighbour rule: (dx*dx+dy*dy)+dz*dz <= rc*rc in
 *   float64 without fused multiply-add. Rows with more than 2048 neighbours
 *   return INCH_ERR_CAPACITY. */
inch_status inch_neighbors_celllist(int64_t n, const double *xyz, double rc, int64_t *indptr,
                                    int32_t *indices, int64_t capacity, int64_t *nnz_out,
                                    void *stream);
/* The neighbour CSR renumbered by a permutation (no reference counterpart as one call: the
 * reference searches the reordered coordinates again, Fuel/Coordinate/T1.py:331-481 after
 * :171-323). order[i] = old index of new atom i (int32), undo[old] = new index (int64);
 * row i of the output is row order[i] of the input with every column j replaced by undo[j],
 * ascending: the CSR a fresh search of the reordered coordinates returns. out_indptr [n+1],
 * out_indices [nnz]. Rows over 2048 entries return INCH_ERR_CAPACITY. Syncs the stream. */
inch_status inch_csr_permute(int64_t n, const int64_t *indptr, const int32_t *indices, const int32_t *order,
                             const int64_t *undo, int64_t *out_indptr, int32_t *out_indices, void *stream);

/* ANM Hessian A = H + plus_i*I into full-storage block rows on the pattern
 * (rowptr, col) (the neighbour CSR). vals64 (float64[nnzb*9]) and/or vals32
 * (float32[nnzb*9], round-to-nearest of the float64 value) may be NULL.
 * diag (float64[3n], optional) receives the scalar diagonal A_diag. */
inch_status inch_hessian_assemble(int64_t n, const double *xyz, double rc, double plus_i,
                                  const int64_t *rowptr, const int32_t *col, double *vals64,
                                  float *vals32, double *diag, void *stream);

/* Block SpMV with fused epilogue (all vectors length 3n, same dtype vdt):
 *   w   = A x                                   (A in dtype mdt, accumulate in mdt)
 *   out = s1 * (w - c0 * x) - c2 * z            (z may be NULL => no z term)
 *   if qcoef: out += sum_m qcoef[m] * Q[m, :]   (Q float64[m_q, 3n], qcoef float64 device)
 *   y   = out;  if acc: acc += mu * out
 * Plain y = A x is s1 = 1, c0 = 0, z = acc = qcoef = NULL. */
inch_status inch_bsr_spmv(int64_t n, const int64_t *rowptr, const int32_t *col, const void *vals,
                          int32_t mdt, int32_t vdt, const void *x, void *y, double s1, double c0,
                          const void *z, double c2, void *acc, double mu, const double *Q,
                          const double *qcoef, int32_t m_q, int32_t lanes_per_row, void *stream);

/* Reference half-storage views of the same block matrix: part 0 = A, 1 = L (scalar
 * lower triangle incl. diagonal, what the reference's CSR `A @ v` computes), 2 = L^T
 * (`A.T @ v`). Compatibility path, not tuned. */
inch_status inch_bsr_spmv_part(int64_t n, const int64_t *rowptr, const int32_t *col, const void *vals,
                               int32_t mdt, int32_t vdt, const void *x, void *y, int32_t part, void *stream);

/* Several right-hand sides: Y[:, r] = A X[:, r] for r < nrhs, X/Y row-major
 * [3n, nrhs] (the nrhs values of one dof are contiguous). */
inch_status inch_bsr_spmm(int64_t n, const int64_t *rowptr, const int32_t *col, const void *vals,
                          int32_t mdt, int32_t vdt, const void *X, void *Y, int32_t nrhs,
                          void *stream);

/* Matrix-free FP64 ANM operator recomputed from float64 coordinates and the
 * neighbour CSR: y = (H + plus_i I) x. Used for the FP64 residual polish. */
inch_status inch_anm_matfree_apply(int64_t n, const double *xyz, const int64_t *rowptr,
                                   const int32_t *col, double plus_i, const double *x, double *y,
                                   int32_t nrhs, void *stream);

/* Matrix-free ANM product with the fused epilogue of inch_bsr_spmv (vectors float32
 * or float64 per vdt; geometry and accumulation in float64). config = lanes*10 + unroll
 * (lanes per row 4/8/16/32, neighbours per lane iteration 1/2; 0 = 161), the autotuned
 * launch shape; + 2000 maps each XCD's workgroups to one contiguous range of rows. */
inch_status inch_anm_matfree_spmv(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                                  double plus_i, int32_t vdt, const void *x, void *y, double s1, double c0,
                                  const void *z, double c2, void *acc, double mu, const double *Q,
                                  const double *qcoef, int32_t m_q, int32_t config, void *stream);

/* Periodic images in x and y (membrane boxes; replaces X_cKDTreePbcXy,
 * Fuel/Coordinate/T1.py:67-161, and the image terms of the Hessian builder,
 * Fuel/T1.py:98-128,317-396). Image s = (sx, sy) in {-1,0,1}^2 shifts atom j by
 * (sx*box_x, sy*box_y, 0); d_s = (x_j + s box) - x_i. An entry (i, j) of the neighbour CSR
 * exists when any image is within the cutoff; images[e] (uint16) holds the mask of those
 * images, bit (sx+1)*3 + (sy+1). box_x, box_y > rc (else "ABORTED." as the reference);
 * (origin_x, origin_y) is the box corner used to wrap coordinates into periodic cells.
 * Two-pass calling convention as inch_neighbors_celllist (images NULL in the count pass). */
inch_status inch_neighbors_celllist_pbc(int64_t n, const double *xyz, double rc, double box_x, double box_y,
                                        double origin_x, double origin_y, int64_t *indptr, int32_t *indices,
                                        uint16_t *images, int64_t capacity, int64_t *nnz_out, void *stream);
/* H_ij = sum over the images in the mask of -d_s d_s^T / |d_s|^2; diagonal as the open case. */
inch_status inch_hessian_assemble_pbc(int64_t n, const double *xyz, double plus_i, const int64_t *rowptr,
                                      const int32_t *col, const uint16_t *images, double box_x, double box_y,
                                      double *vals64, float *vals32, double *diag, void *stream);
/* Matrix-free product with images: y = (H + plus_i I) x for nrhs row-major right-hand sides
 * (x, y [3n, nrhs]); with nrhs == 1 the epilogue of inch_anm_matfree_spmv applies. */
inch_status inch_anm_matfree_pbc(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                                 const uint16_t *images, double box_x, double box_y, double plus_i, int32_t vdt,
                                 const void *x, void *y, int32_t nrhs, double s1, double c0, const void *z,
                                 double c2, void *acc, double mu, const double *Q, const double *qcoef, int32_t m_q,
                                 void *stream);

/* inch_anm_matfree_spmv with 16-bit column offsets (rowbase / col16 from inch_bsr_col16). */
inch_status inch_anm_matfree_spmv16(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *rowbase,
                                    const uint16_t *col16, double plus_i, int32_t vdt, const void *x, void *y,
                                    double s1, double c0, const void *z, double c2, void *acc, double mu,
                                    const double *Q, const double *qcoef, int32_t m_q, int32_t config, void *stream);

/* inch_bsr_spmv with 16-bit column offsets (same reference interfaces): column j of block b
 * in row i is rowbase[i] + col16[b]; inch_bsr_col16 builds both from the int32 columns and
 * fails with INCH_ERR_CAPACITY when a row spans 65,536 columns or more (syncs).
 * lanes_per_row + 1000 maps each XCD's workgroups to one contiguous range of rows. */
inch_status inch_bsr_spmv16(int64_t n, const int64_t *rowptr, const int32_t *rowbase, const uint16_t *col16,
                            const void *vals, int32_t mdt, int32_t vdt, const void *x, void *y, double s1, double c0,
                            const void *z, double c2, void *acc, double mu, const double *Q, const double *qcoef,
                            int32_t m_q, int32_t lanes_per_row, void *stream);
inch_status inch_bsr_col16(int64_t n, const int64_t *rowptr, const int32_t *col, int32_t *rowbase,
                           uint16_t *col16, void *stream);

/* dst = src (bytes, multiple of 16): streaming copy for measuring achievable HBM bandwidth. */
inch_status inch_stream_copy(int64_t bytes, const void *src, void *dst, void *stream);

/* Cluster layout of the matrix-free operator: another layout of the neighbour CSR that
 * inch_anm_matfree_spmv reads (same reference interfaces). The atoms are grouped into
 * ncl = ceil(n / 4) clusters that are compact in space: cells of cell_xy x cell_xy x cell_z
 * (z cells coarsened until there are at most max(16n, 65536) cells), numbered by x, y column
 * then z, atoms in a cell by index, cut into groups of 4; the groups are then numbered by
 * their smallest neighbour index (ties by spatial order), so consecutive clusters gather
 * from one band of the vectors; cluster_atoms int32[ncl * 4] (-1 pads the last group). Each cluster lists the union of its
 * atoms' rows once: entries uint32 = ((j - base[c]) << 8) | mask, ascending j, mask bit t set
 * when j is a neighbour of atom t of the cluster at a nonzero distance (j itself and
 * coincident atoms, whose terms the product masks, excluded); cl_ptr int64[ncl + 1],
 * base int32[ncl]. Two passes, as inch_neighbors_celllist: with entries == NULL it writes
 * cluster_atoms, cl_ptr, base and *nnz_out; the call with entries (capacity >= nnz) and the
 * arrays of the first pass writes the entries. INCH_ERR_CAPACITY when a cluster's rows hold
 * more than 4096 entries or its neighbours span 2^24 columns or more (syncs). */
inch_status inch_anm_clusters(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                              double cell_xy, double cell_z, int32_t *cluster_atoms, int64_t *cl_ptr,
                              int32_t *base, uint32_t *entries, int64_t capacity, int64_t *nnz_out, void *stream);
/* The same layout in one call: entries must have room for the layout (capacity >= its nnz;
 * the neighbour CSR's entry count rowptr[n] always suffices). The count pass writes every
 * cluster's list into a provisional buffer (rowptr[n] words, allocated and freed here) and
 * the lists are copied into the final numbering, instead of a second merge of the rows.
 * Same outputs, bit for bit, as the two calls; the same errors. Syncs. */
inch_status inch_anm_clusters_one_pass(int64_t n, const double *xyz, const int64_t *rowptr, const int32_t *col,
                                       double cell_xy, double cell_z, int32_t *cluster_atoms, int64_t *cl_ptr,
                                       int32_t *base, uint32_t *entries, int64_t capacity, int64_t *nnz_out,
                                       void *stream);
/* inch_anm_matfree_spmv on the cluster layout: the same product and fused epilogue, the
 * neighbour's coordinates and vector entries loaded once per cluster (one wave per cluster,
 * lanes over its list). config = list entries per lane iteration (1, 2 or 4; 0 = 2002), or
 * 100 + clusters per wave (4, 8 or 16) for the persistent form, which prefetches the next
 * cluster's list while computing the current one (bitwise-equal results); + 2000 maps each
 * XCD's workgroups to one contiguous range of clusters. */
inch_status inch_anm_matfree_cl(int64_t n, const double *xyz, int64_t ncl, const int32_t *cluster_atoms,
                                const int64_t *cl_ptr, const int32_t *base, const uint32_t *entries, double plus_i,
                                int32_t vdt, const void *x, void *y, double s1, double c0, const void *z, double c2,
                                void *acc, double mu, const double *Q, const double *qcoef, int32_t m_q,
                                int32_t config, void *stream);
/* inch_anm_matfree_apply on the cluster layout: Y = (H + plus_i I) X for nrhs row-major
 * right-hand sides (X, Y float64 [3n, nrhs]). */
inch_status inch_anm_matfree_cl_apply(int64_t n, const double *xyz, int64_t ncl, const int32_t *cluster_atoms,
                                      const int64_t *cl_ptr, const int32_t *base, const uint32_t *entries,
                                      double plus_i, const double *x, double *y, int32_t nrhs, void *stream);
/* Record form of the cluster layout for the Chebyshev filter's recurrence (replaces the
 * recurrence vectors of ChebAv, Burn/PolynomialFilters/T2.py:28-92, around the product of
 * Burn/Krylov/T3.py:54-239). Slot q = 4 c + t is atom t of cluster c. inch_anm_records
 * writes rec float64 [24 ncl] (slot q: coordinates at 6q..6q+2, zeros for a padding slot;
 * vector part 6q+3..6q+5 zeroed) and rec_entries uint32 [nnz of the layout] = 48 (slot of
 * j) | mask per cluster (the byte offset of j's record, low 4 bits holding the mask),
 * ascending slot (cl_ptr unchanged). INCH_ERR_CAPACITY when
 * 192 ncl exceeds 2^32 - 1 (32-bit record offsets). Syncs. */
inch_status inch_anm_records(int64_t n, const double *xyz, int64_t ncl, const int32_t *cluster_atoms,
                             const int64_t *cl_ptr, const int32_t *base, const uint32_t *entries, double *rec,
                             uint32_t *rec_entries, void *stream);
/* One step of the recurrence on records: result = s1 ((H + plus_i I) x - c0 x) - c2 z, with
 * x, z the vector parts of rec_in, rec_z (rec_z may be NULL: no z term), written to the vector
 * part of rec_out and/or to y (float64 [12 ncl], slot order; either may be NULL, not both);
 * acc float64 [12 ncl] (slot order, may be NULL) += mu times the result. rec_out's coordinate
 * part is not written; rec_in's coordinates are read. xcd != 0 maps each XCD's workgroups to
 * one contiguous range of clusters. */
inch_status inch_anm_matfree_rec(int64_t ncl, const double *rec_in, const double *rec_z, double *rec_out, double *y,
                                 const int64_t *cl_ptr, const uint32_t *rec_entries, double plus_i, double s1,
                                 double c0, double c2, double *acc, double mu, int32_t xcd, void *stream);
/* Vector part of rec for slot q = v[3 a .. 3 a + 2] with a = cluster_atoms[q] (0 for padding),
 * or v[3 q .. 3 q + 2] when cluster_atoms is NULL (v in slot order, [12 ncl]); acc (may be
 * NULL) [12 ncl] = mu0 times the same. */
inch_status inch_rec_pack(int64_t ncl, const int32_t *cluster_atoms, const double *v, double *rec, double *acc,
                          double mu0, void *stream);
/* u[3 a .. 3 a + 2] = acc of slot q for every non-padding slot q (a = cluster_atoms[q]). */
inch_status inch_rec_unpack(int64_t ncl, const int32_t *cluster_atoms, const double *acc, double *u, void *stream);

/* h[j] = sum_i V[j*ld + i] * u[i] for j < J (float64, deterministic two-phase
 * reduction). work: float64 scratch of inch_multidot_work_size(J) elements. */
int64_t inch_multidot_work_size(int32_t J);
inch_status inch_multidot(int64_t len, int32_t J, const double *V, int64_t ld, const double *u,
                          double *h, double *work, void *stream);

/* u[i] -= scale * sum_j V[j*ld + i] * h[j]  (h on device).
 * If nrm2_out: *nrm2_out = ||u_new||_2 (two-phase, work as in inch_multidot). */
inch_status inch_gemv_update(int64_t len, int32_t J, const double *V, int64_t ld, const double *h,
                             double scale, double *u, double *nrm2_out, double *work, void *stream);

/* dst[i] = src[i] / (*div)   (div: device scalar), optionally also into dst2. */
inch_status inch_vec_scale_div(int64_t len, const double *src, const double *div, double *dst,
                               double *dst2, void *stream);
/* y[i] += a[0] * x[i] * sign  (a: device scalar). */
inch_status inch_vec_axpy_dev(int64_t len, const double *a, double sign, const double *x, double *y,
                              void *stream);
/* y = a*x + b*y elementwise (any float32/float64 combination; b == 0 ignores y). */
inch_status inch_vec_axpby(int64_t len, double a, const void *x, int32_t xdt, double b, void *y, int32_t ydt,
                           void *stream);
/* Type conversion between float64 and float32 vectors. */
inch_status inch_vec_convert(int64_t len, const void *src, int32_t sdt, void *dst, int32_t ddt,
                             void *stream);

/* G[a*k2 + b] = sum_i X[a*ldx + i] * Y[b*ldy + i]  (a < k1, b < k2; float64 MFMA
 * 16x16x4 tiles, deterministic two-phase reduction). work: inch_gram_work_size(k1, k2)
 * doubles. */
int64_t inch_gram_work_size(int32_t k1, int32_t k2);
inch_status inch_gram(int64_t len, int32_t k1, const double *X, int64_t ldx, int32_t k2, const double *Y,
                      int64_t ldy, double *G, double *work, void *stream);

/* out[j] = || Y[j*ldy + :] - theta[j] * X[j*ldx + :] ||_2 for j < k (work >= 256*k doubles). */
inch_status inch_residual_norms(int64_t len, int32_t k, const double *X, int64_t ldx, const double *Y, int64_t ldy,
                                const double *theta, double *out, double *work, void *stream);

/* One fused CGS2 re-orthogonalisation of a Lanczos step against the basis V[:J]
 * (J <= 256), given h1 = V[:J] u from inch_multidot:
 *   u' = u - V^T h1;  beta = ||u'||;  h2 = V[:J] u';  u = u' - V^T h2;  vnext = u / beta
 * (vnext may be NULL). beta_out: device scalar. work: inch_cgs2_work_size(J) doubles.
 * Reads V twice (three times with the first projection) instead of four. */
int64_t inch_cgs2_work_size(int32_t J);
inch_status inch_cgs2_reorth(int64_t len, int32_t J, const double *V, int64_t ld, const double *h1, double *u,
                             double *vnext, double *beta_out, double *work, void *stream);
/* Lanczos orthogonalisation with delayed reorthogonalisation (DCGS2; no reference
 * counterpart as one call: it replaces the CGS2 sequence of inch_multidot +
 * inch_cgs2_reorth, TRLM/T1.py:465-520, with two reads of V instead of three). i = J - 1;
 * V[i] is the provisional vector of the previous step and u = Op V[i]. One read of V[0..i]
 * gives s = V[0..i] u and g = V[0..i-1] V[i]; a second read sets V[i] -= V[0..i-1] g and
 * projects u off the corrected basis (u - sum_{j<i} s_j V_j - (s_i - g.s) v_i).
 * alpha[i] = s_i - g.s - beta[i-1] g[i-1] (the only operator term that survives in a Krylov
 * basis, thick restarts included). Writes alpha[i], *beta_out = |u| and vnext = u / |u|
 * (when non-NULL). work: inch_dcgs2_work_size(J) doubles. */
int64_t inch_dcgs2_work_size(int32_t J);
inch_status inch_dcgs2_step(int64_t len, int32_t J, double *V, int64_t ld, double *u, double *alpha,
                            const double *beta, double *beta_out, double *vnext, double *work, void *stream);
/* Allocation-free form of inch_dcgs2_step: the same kernels and results, with the step's
 * scratch in a static device array of the library instead of `work` (J <= 256; larger J
 * returns INCH_ERR_CAPACITY). No counterpart in the reference: it is for callers that cannot
 * allocate device memory (HIP graph capture, serving loops). It allocates nothing; the array
 * (2.0 MiB per GPU) is held until the process ends, and its device address is looked up once
 * per device on the first call there. Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other host threads) race. No word of the array
 * needs clearing between calls. */
inch_status inch_dcgs2_step_noalloc(int64_t len, int32_t J, double *V, int64_t ld, double *u, double *alpha,
                                    const double *beta, double *beta_out, double *vnext, void *stream);

/* out[q*ld_out + i] = sum_r W[r*k + q] * V[r*ld + i]  for q < k, r < m
 * (W float64 [m, k] row-major on device; out must not alias V). FP64 MFMA 16x16x4. */
inch_status inch_rr_combine(int64_t len, int32_t m, int32_t k, const double *V, int64_t ld,
                            const double *W, double *out, int64_t ld_out, void *stream);

/* ------------------------------------- learned warm start (graph network) ---- */
/* No counterpart in the reference (its solvers start from random vectors). Layer widths
 * <= 96; X/W/Y float16 row-major, accumulation and gradients float32. */
/* Y = act(X W + b) for X [N, Din], W [Din, Dout]; Ypre (float32) and Y32 optional. */
inch_status inch_gnn_dense_fwd(int64_t N, int32_t Din, int32_t Dout, const void *X, const void *W, const float *b,
                               int32_t relu, float *Ypre, void *Y, float *Y32, void *stream);
/* dX = dY W^T (dX may be NULL), dW = X^T dY, db = sum_i dY[i]; work: inch_gnn_work_size floats. */
int64_t inch_gnn_work_size(int32_t Din, int32_t Dout);
inch_status inch_gnn_dense_bwd(int64_t N, int32_t Din, int32_t Dout, const void *X, const void *W, const float *dY,
                               float *dX, float *dW, float *db, float *work, void *stream);
/* Neighbour mean over the ANM graph (self excluded): mode 0 Y[i] = mean_{j in N(i)} X[j];
 * mode 1 its adjoint. dtype codes: 0 float32, 2 float16. */
inch_status inch_gnn_aggregate(int64_t N, int32_t D, const int64_t *rowptr, const int32_t *col, const void *X,
                               int32_t xdt, void *Y, int32_t ydt, int32_t mode, void *stream);
/* dPre = dY where pre > 0 else 0. */
inch_status inch_gnn_relu_bwd(int64_t n, const float *pre, const float *dY, float *dPre, void *stream);

/* ---------------------------------------------- batched (segmented) FP64 ---- */
/* Many structures stored as contiguous dof segments of one vector. Chunk c covers dofs
 * [cbeg[c], cend[c]) of segment cseg[c]; segment s owns chunks sptr[s] .. sptr[s+1]-1.
 * Replaces, for a batch of structures at once, the per-structure cuBLAS dot/nrm2/gemv
 * calls of Burn/ThickRestartLanczosHotellingDeflation/T1.py:465-520 and the restart GEMM
 * of :176,246 (the reference solves one structure per driver loop iteration). */
/* h[s*J + j] = sum_{i in s} V[j*ld + i] u[i]; work >= nchunk*J doubles. */
inch_status inch_seg_multidot(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend,
                              const int64_t *sptr, int32_t J, const double *V, int64_t ld, const double *u, double *h,
                              double *work, void *stream);
/* u[i] -= scale * sum_j V[j*ld+i] h[s*hld+j] (J <= 512); nrm_out[s] = ||u_s|| if given
 * (work >= nchunk doubles). */
inch_status inch_seg_update(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg,
                            const int64_t *sptr, int32_t J, const double *V, int64_t ld, const double *h, int32_t hld,
                            double scale, double *u, double *nrm_out, double *work, void *stream);
/* Rigid-body (Hotelling) projection per segment from the coordinates (replaces the products
 * with the six orthonormal rigid-body rows of the driver's deflation basis,
 * commandBenchmark_InchingTRLMHD...:314-372, used by
 * Burn/ThickRestartLanczosHotellingDeflation/T1.py:465-520): y -= scale * P x, where at atom i
 * of segment s, P x = mean_s(x) + (K_s^{-1} sum_j r_j x x_j) x r_i with r_i = xyz_i - cen[s],
 * K_s = sum_i (|r_i|^2 I - r_i r_i^T). xyz float64 [n, 3] in the vector's atom order (dof 3i+a
 * is component a of atom i), cen [S, 3] the segments' centroids, kinv [S, 9] row-major
 * K_s^{-1}, ninv [S] = 1 / atoms of s. x and y [3n] may be the same vector (y <- y - scale P y).
 * g receives the S x 6 moments (sum of x's components, sum r_i x x_i); with y NULL only they
 * are formed (kinv and ninv may then be NULL). work >= nchunk * 6 doubles. */
inch_status inch_seg_rigid_apply(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend,
                                 const int32_t *cseg, const int64_t *sptr, const double *xyz, const double *cen,
                                 const double *kinv, const double *ninv, const double *x, double *y, double scale,
                                 double *g, double *work, void *stream);
/* CGS2's update and second projection in one read of V, per segment: u -= V[:J] h (h of
 * each segment, row stride hld), then h2[s, j] = V[j] . u over segment s (j < J),
 * h2[s, J] = |u|^2 over s and nrm_out[s] = sqrt(h2[s, J]); h2 is [S, J + 1]. work holds
 * nchunk * (J + 1) partials. J <= 256. */
inch_status inch_seg_cgs_b(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg,
                           const int64_t *sptr, int32_t J, const double *V, int64_t ld, const double *h, int32_t hld,
                           double *u, double *h2, double *nrm_out, double *work, void *stream);
/* Delayed CGS2 Lanczos step per segment (the segmented inch_dcgs2_step; replaces the
 * per-structure CGS2 of TRLM/T1.py:465-520 with two reads of V instead of three): i = J - 1,
 * V[i] provisional, u = Op V[i]. Corrects V[i] -= V[:i] g per segment, projects u off the
 * corrected V[:J], alpha_i[s] = v_i . u - beta_prev[s] g_{i-1}, beta_out[s] = |u_s|.
 * work: inch_seg_dcgs2_work_size doubles. J <= 256. */
int64_t inch_seg_dcgs2_work_size(int32_t nchunk, int32_t S, int32_t J);
inch_status inch_seg_dcgs2_step(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend,
                                const int32_t *cseg, const int64_t *sptr, int32_t J, double *V, int64_t ld, double *u,
                                double *alpha_i, const double *beta_prev, double *beta_out, double *work,
                                void *stream);
/* dst[i] = src[i] / div[s] (and dst2). */
inch_status inch_seg_scale(int32_t nchunk, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg,
                           const double *src, const double *div, double *dst, double *dst2, void *stream);
/* out[q*ld_out+i] = sum_r W[s][r][q] V[r*ld+i] (W [S, m, k], k <= 64; W[s] padded to
 * [4 ceil(m/4)][16 ceil(k/16) + 1] doubles must fit 64 KiB of LDS). FP64 MFMA 16x16x4. */
inch_status inch_seg_combine(int32_t nchunk, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg, int32_t m,
                             int32_t k, const double *V, int64_t ld, const double *W, double *out, int64_t ld_out,
                             void *stream);
/* out[s*k+q] = || Y[q] - theta[s*k+q] X[q] || over segment s; work >= nchunk*k doubles. */
inch_status inch_seg_resnorm(int32_t nchunk, int32_t S, const int64_t *cbeg, const int64_t *cend, const int32_t *cseg,
                             const int64_t *sptr, int32_t k, const double *X, int64_t ldx, const double *Y,
                             int64_t ldy, const double *theta, double *out, double *work, void *stream);

/* ------------------------------------------------- analysis of computed modes ---- */
/* Square fluctuations of the modes eigvec [k, N, 3] (float64, atom order of the caller):
 * ratio_var[m] = r_m / max r, r_m = 1 / sqrt(|eigval[m]|); sfpm[m, a] = |eigvec[m, a, :]|;
 * ratio[m, a] = ratio_var[m] sfpm[m, a]; sqfluct[a] = sum_m ratio[m, a]. eigval is a device
 * pointer. Replaces Heigval_Heigvec_SqFluctPerMode_SqFluct_RatioSqFPM_RatioVar,
 * Burn/Visualisation/T2.py:43-55. */
inch_status inch_sqfluct(int32_t k, int64_t N, const double *eigval, const double *eigvec, double *ratio_var,
                         double *sfpm, double *ratio, double *sqfluct, void *stream);
/* Cross-correlation maps of P pairs of modes: out[p, a, b] (float32 [P, N, N]) =
 * sign(s) log2(|s| + 1), s = float32((eigvec[i, a, :] . eigvec[j, b, :]) eigval[i] eigval[j])
 * * 100^precision, (i, j) = pairs[2p], pairs[2p+1]. pairs (int32 [P, 2]) and eigval
 * (float64 [k]) are HOST pointers; every index is checked against k; precision in [0, 4].
 * Replaces the hessian_CC loop and S_LogModulusS of Heigval_Heigvec_HccPairlist_HccBatch,
 * Burn/Visualisation/T2.py:58-116 and Burn/Visualisation/T1.py:44-52. */
inch_status inch_cross_correlation(int32_t P, const int32_t *pairs, int32_t k, int64_t N, const double *eigval,
                                   const double *eigvec, int32_t precision, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* INCHING_H */
