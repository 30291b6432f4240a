"""CPU baselines of the time-to-k-modes workloads, timed with the oracle (numpy/scipy).

``cpu_baseline_assembly`` (the disconnected assembly) is described in its docstring; the
rest of this header describes ``cpu_baseline`` (one connected globule).

The reference's CPU-side path is its own Python/numpy code (3DRCM loop over k-d tree
queries, ``InchingLiteInteger/Fuel/Coordinate/T1.py:171-323``) plus an FP64 L+D CSR
operator applied as L v + L^T v - diag v (``Burn/Krylov/T3.py:131-160``) inside TRLM with
CGS2 re-orthogonalisation (``Burn/ThickRestartLanczosHotellingDeflation/T1.py:415-522``).
A full CPU solve of the 300k-atom workload takes hours, so a bounded sample is timed and
scaled to the operation counts of the GPU run:

* set-up (neighbours + RCM + CSR Hessian) on a ``sample_atoms`` globule of the same
  density, scaled linearly in the atom count;
* one L+D CSR matvec on that sample, scaled by the block-count ratio, times the GPU run's
  SpMV count;
* one CGS2 re-orthogonalisation step at the full size (numpy BLAS, all cores) with the
  average basis length, times the GPU run's Lanczos step count.
"""
import os
import time

import numpy as np

from . import geometry as OG
from . import hessian as OH
from . import rcm as OR


def cpu_baseline(X, full_atoms, full_nnzb, k, gpu_result):
    """X: coordinates [sample_atoms, 3] of a globule with the workload's density."""
    cores = len(os.sched_getaffinity(0))
    X = np.asarray(X, dtype=np.float64)
    sample_atoms = X.shape[0]
    t = time.time()
    nb = OG.neighbour_csr(X, 8.0)
    t_nb = time.time() - t
    t = time.time()
    order, _ = OR.rcm_order(*nb)
    t_rcm = time.time() - t
    t = time.time()
    L, d = OH.lower_csr_and_diag(X[order], 8.0, 1.0)
    t_h = time.time() - t
    sample_nnzb = nb[1].size
    v = np.random.default_rng(0).standard_normal(3 * sample_atoms)
    OH.halfmem_matvec(L, d, v)
    reps = 5
    t = time.time()
    for _ in range(reps):
        OH.halfmem_matvec(L, d, v)
    t_spmv = (time.time() - t) / reps
    n = 3 * full_atoms
    ncv = int(getattr(gpu_result, "ncv", 2 * k))
    j = (ncv + k) // 2
    V = np.random.default_rng(1).standard_normal((j, n))
    u = np.random.default_rng(2).standard_normal(n)
    t = time.time()
    for _ in range(2):
        for _p in range(2):
            h = V @ u
            u -= V.T @ h
    t_cgs = (time.time() - t) / 2
    del V
    scale_n = full_atoms / sample_atoms
    scale_b = full_nnzb / max(1, sample_nnzb)
    setup = (t_nb + t_rcm + t_h) * scale_n
    spmv = gpu_result.total_spmv * t_spmv * scale_b
    cgs = gpu_result.n_apply * t_cgs
    total = setup + spmv + cgs
    return {"value": round(total, 2), "unit": "s", "cores": cores, "kind": "port",
            "sample": ("oracle on a %d-atom globule (neighbours %.2fs, RCM %.2fs, CSR Hessian %.2fs, "
                       "L+D matvec %.4fs) scaled x%.1f atoms / x%.1f blocks to %d matvecs; numpy CGS2 step at "
                       "n=%d, j=%d: %.4fs x %d Lanczos steps; scipy matvec single-threaded, BLAS on %d cores")
            % (sample_atoms, t_nb, t_rcm, t_h, t_spmv, scale_n, scale_b, gpu_result.total_spmv, n, j, t_cgs,
               gpu_result.n_apply, cores),
            "breakdown_s": {"setup": round(setup, 2), "spmv": round(spmv, 2), "reorth": round(cgs, 2)}}


def cpu_baseline_assembly(X_sub, n_components, k_with_rigid, tol_rel, lam_max, theta_max, rc=8.0, plus_i=1.0):
    """CPU time of the disconnected-assembly workload by the reference's CPU method, timed on
    one of its subunits and multiplied by the number of subunits.

    The assembly's Hessian is block-diagonal with one block per subunit, so its lowest modes
    are the union of the subunits' lowest modes and a CPU solve of the assembly is a solve
    of each subunit. The sample is one subunit ``X_sub`` (coordinates [N_sub, 3]; a rotation
    does not change its spectrum), taken end to end: neighbours, 3DRCM and the CSR Hessian
    with the oracle (``Fuel/Coordinate/T1.py:171-323``, ``Fuel/T1.py:228-660``), then
    ``scipy.sparse.linalg.eigsh(H + PlusI I, k, which='SA')``, the ARPACK call of the
    reference's CPU baseline (``Command8A/commandArpack_Pdb_0064_EPYC.py:328``), with
    ``k_with_rigid`` modes (the subunit's wanted modes and its six rigid-body modes).
    ARPACK stops at ||r|| <= tol |theta|; tol = tol_rel lam_max / theta_max makes that the
    relative-residual target of the GPU run (``theta_max``: its largest wanted eigenvalue).
    The residuals of the CPU result are checked against tol_rel afterwards."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    cores = len(os.sched_getaffinity(0))
    X = np.asarray(X_sub, dtype=np.float64)
    N = X.shape[0]
    t0 = time.time()
    ip, ix = OG.neighbour_csr(X, rc)
    t_nb = time.time() - t0
    t = time.time()
    order, _ = OR.rcm_order(ip, ix)
    t_rcm = time.time() - t
    t = time.time()
    Xp = X[order]
    rp, col, v = OH.block_hessian(Xp, rc, plus_i)
    H = sp.bsr_matrix((v, col.astype(np.int64), rp), shape=(3 * N, 3 * N)).tocsr()
    t_h = time.time() - t
    tol = tol_rel * lam_max / theta_max
    t = time.time()
    w, U = sla.eigsh(H, k=k_with_rigid, which="SA", tol=tol)
    t_eig = time.time() - t
    rel = np.linalg.norm(H @ U - U * w[None, :], axis=0) / lam_max
    per_sub = time.time() - t0
    return {"value": round(per_sub * n_components, 1), "unit": "s", "cores": cores, "kind": "port",
            "sample": ("one of the %d subunits (%d atoms, the first subunit's globule) solved completely on the "
                       "host: oracle neighbours %.2fs + 3DRCM %.2fs + CSR Hessian %.2fs, then scipy eigsh(k=%d, "
                       "which='SA', tol=%.2e) %.2fs (ARPACK, the reference's CPU method; matvec single-threaded, "
                       "BLAS on %d cores); max rel_res %.1e; value = %d x %.2fs")
            % (n_components, N, t_nb, t_rcm, t_h, k_with_rigid, tol, t_eig, cores, float(rel.max()),
               n_components, per_sub),
            "per_subunit_s": round(per_sub, 3), "max_rel_res": float(rel.max()),
            "lowest_nontrivial": [float(x) for x in np.sort(w)[6:9]]}
