"""Workload configurations at their stated sizes, the output convention on device results,
and the failure paths of the solvers.

* The plain method on a 5,000-atom globule (seed 0, 64 modes, tol 1e-9), float64 stored and
  matrix-free operators, against dense eigh of the 15,000 x 15,000 oracle Hessian.
* Modes returned in the original atom order (driver ``:396-407``) checked against a Hessian
  built from the coordinates in their original order.
* The learned warm start on unseen globules: a small one against dense eigh (TRLM and JDM),
  and an unseen 300,000-atom globule through size-independent properties.
* Failure paths: maxiter exhausted (warning, flag, best pairs), an undersized HalfNnz buffer
  (IndexError, ``Fuel/T1.py:486-536``), non-finite coordinates (INCH_ERR_NAN), and a warm
  block from a corrupted network (cold-start fallback).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import hessian as OH
from tests.helpers import small_globule

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(__file__)), "inching_amd", "weights", "warmstart_gnn.npz")


def _oracle_csr(Xp, rc=8.0, plus_i=1.0):
    rp, col, v = OH.block_hessian(Xp, rc, plus_i)
    n = 3 * Xp.shape[0]
    return sp.bsr_matrix((v, col.astype(np.int64), rp), shape=(n, n)).tocsr()


def _rel_res_cpu(H, lam, Vrows, lam_max):
    """Relative residuals ||H v - lam v|| / (lam_max ||v||) with the oracle's CSR Hessian."""
    V = Vrows.T
    R = H @ V - V * lam[None, :]
    return np.linalg.norm(R, axis=0) / (lam_max * np.linalg.norm(V, axis=0))


@pytest.fixture(scope="module")
def plain5000(cuda):
    """The 5,000-atom globule (seed 0) in RCM order, its oracle CSR Hessian and the dense
    eigenvalues (eigvalsh of 15,000 x 15,000, computed once)."""
    from inching_amd import synth
    from inching_amd.hessian import build_hessian
    X = synth.globule(5000, seed=0)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=("f64",))
    H = _oracle_csr(X[A.order])
    print("case: dense eigvalsh 15000", flush=True)
    return A, H, np.linalg.eigvalsh(H.toarray())


@pytest.mark.parametrize("precision", ["f64", "mf64"])
def test_plain_method_5000_atoms_64_modes(plain5000, precision):
    """5,000-atom globule (seed 0), 64 non-trivial modes at tol 1e-9: eigenvalues against
    dense eigvalsh of the oracle's Hessian, residuals recomputed on the host with the
    oracle's CSR matrix."""
    from inching_amd.hessian import rigid_body_basis
    from inching_amd.pipeline import lowest_modes
    print("case: 5000 atoms k=64 tol 1e-9 %s" % precision, flush=True)
    A, H, w = plain5000
    Q = rigid_body_basis(A.xyz)
    r = lowest_modes(A, 64, Q=Q, precision=precision, tol_rel=1e-9)
    assert r.converged and np.max(r.rel_res) <= 1e-9
    assert np.max(np.abs(r.eigval - w[6:70])) <= 1e-9, np.max(np.abs(r.eigval - w[6:70]))
    rel = _rel_res_cpu(H, r.eigval, r.eigvec.cpu().numpy(), w[-1])
    assert rel.max() <= 1.5e-9, rel.max()


def test_modes_in_atom_order_on_device_results(cuda):
    """The (k, N, 3) modes in the original atom order are eigenvectors of the Hessian built
    from the original (unpermuted) coordinates, and the eigenvalues are those of H."""
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes
    from inching_amd.structure import modes_in_atom_order
    print("case: modes in atom order", flush=True)
    X = small_globule(1500, seed=61)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
    assert not np.array_equal(A.order, np.arange(1500))
    r = lowest_modes(A, 10, Q=rigid_body_basis(A.xyz), precision="mf64", tol_rel=1e-10)
    lam_h, modes = modes_in_atom_order(torch.as_tensor(r.eigval, device="cuda"), r.eigvec, A.undo, plus_i=1.0)
    assert modes.shape == (10, 1500, 3)
    H = _oracle_csr(X, plus_i=0.0)                         # original atom order, H itself
    rel = _rel_res_cpu(H, lam_h, modes.reshape(10, -1), r.lam_max)
    assert rel.max() <= 2e-10, rel.max()
    w = np.linalg.eigvalsh(H.toarray())
    assert np.max(np.abs(lam_h - w[6:16])) <= 1e-9


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained weights not present")
def test_warm_start_unseen_small_globule_matches_dense(cuda):
    """The trained network on an unseen 3,000-atom globule (a seed outside the training set):
    TRLM and Jacobi-Davidson started from its block reach rel_res 1e-8 without falling back,
    and return the eigenvalues of dense eigh."""
    from inching_amd import synth, warmstart as W
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes_jdm_warm, lowest_modes_warm
    print("case: warm start unseen 3000 atoms", flush=True)
    net = W.WarmStartNet.load()
    X = synth.globule(3000, seed=123457)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
    Q = rigid_body_basis(A.xyz)
    w = np.linalg.eigvalsh(_oracle_csr(X[A.order]).toarray())
    k = 24
    for solve in (lowest_modes_warm, lowest_modes_jdm_warm):
        r = solve(A, k, net, Q=Q, precision="mf64", tol_rel=1e-8)
        print("  %s: %d operator launches, fallback %s" % (solve.__name__, r.total_spmv, r.fallback), flush=True)
        assert not r.fallback and r.converged
        assert np.max(r.rel_res) <= 1e-8
        assert np.max(np.abs(r.eigval - w[6:6 + k])) <= 1e-8, np.max(np.abs(r.eigval - w[6:6 + k]))


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained weights not present")
def test_warm_start_unseen_300k_properties(cuda):
    """An unseen 300,000-atom globule (seed 4): the warm TRLM and JDM solves reach rel_res
    1e-8 with no fallback, agree with the cold TRLM eigenvalues, and the warm JDM needs fewer
    operator applications than the cold JDM."""
    from inching_amd import synth, warmstart as W
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes, lowest_modes_jdm, lowest_modes_jdm_warm, lowest_modes_warm
    print("case: warm start unseen 300k atoms", flush=True)
    net = W.WarmStartNet.load()
    X = synth.globule(300000, seed=4)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
    Q = rigid_body_basis(A.xyz)
    k = 32
    ref = lowest_modes(A, k, Q=Q, precision="mf64", tol_rel=1e-8)
    tw = lowest_modes_warm(A, k, net, Q=Q, precision="mf64", tol_rel=1e-8)
    jc = lowest_modes_jdm(A, k, Q=Q, precision="mf64", tol_rel=1e-8)
    jw = lowest_modes_jdm_warm(A, k, net, Q=Q, precision="mf64", tol_rel=1e-8)
    print("  trlm cold %d / warm %d, jdm cold %d / warm %d launches" % (ref.total_spmv, tw.total_spmv,
                                                                        jc.total_spmv, jw.total_spmv), flush=True)
    for r in (ref, tw, jc, jw):
        assert r.converged and np.max(r.rel_res) <= 1e-8
        assert np.max(np.abs(r.eigval - ref.eigval)) <= 1e-8
    assert not tw.fallback and not jw.fallback
    assert jw.total_spmv < jc.total_spmv


def test_maxiter_exhausted_warns_and_flags(cuda):
    """No restart allowed (maxiter=0) and a tolerance out of reach: a NotConvergedWarning,
    converged=False, and finite best pairs from one float64 polish."""
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import NotConvergedWarning, lowest_modes
    print("case: maxiter exhausted", flush=True)
    X = small_globule(800, seed=62)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=("f64",))
    with pytest.warns(NotConvergedWarning, match="ABORTED"):
        r = lowest_modes(A, 12, Q=rigid_body_basis(A.xyz), precision="f64", tol_rel=1e-15, maxiter=0)
    assert r.converged is False
    assert r.eigvec.shape == (12, 2400) and np.all(np.isfinite(r.eigval)) and np.all(np.isfinite(r.rel_res))
    assert np.max(r.rel_res) > 1e-15


def test_halfnnz_buffer_too_small_raises_index_error(cuda):
    """The builder's HalfNnz argument bounds the lower-triangle scalars it may hold, like the
    reference's preallocated buffers (a cupy.put(mode='raise') overflow is an IndexError)."""
    from InchingLiteInteger.Fuel.Coordinate.T1 import X_KdUngappedMinMaxNeighbor
    from InchingLiteInteger.Fuel.T1 import Xnumpy_SparseCupyMatrixUngappped
    print("case: undersized HalfNnz buffer", flush=True)
    X = small_globule(600, seed=63) / 10.0
    isl, half = X_KdUngappedMinMaxNeighbor(X, rc_Gamma=8.0, maxleafsize=100, User_ReturnHalfNnz=True)
    b = Xnumpy_SparseCupyMatrixUngappped(X, maxleafsize=100, rc_Gamma=8.0, User_PlusI=1.0, NnzMinMaxDict=isl)
    A, d = b.ReturnCupyHLowerTriangleInt32(User_MaxHalfNnzBufferSize=half)   # the islands' estimate suffices
    assert A.shape == (1800, 1800)
    # the lower triangle incl. the diagonal holds 9 scalars per off-diagonal block pair and 6
    # per diagonal block (the islands' HalfNnz counts 9 there, a margin of 3 per atom)
    need = (A.nnzb - 600) // 2 * 9 + 600 * 6
    assert need == half - 3 * 600
    b.ReturnCupyHLowerTriangleInt64(User_MaxHalfNnzBufferSize=need)
    with pytest.raises(IndexError):
        b.ReturnCupyHLowerTriangleInt32(User_MaxHalfNnzBufferSize=need - 2)
    with pytest.raises(IndexError):
        b.ReturnCupyHLowerTriangleInt64(User_MaxHalfNnzBufferSize=half / 2)


def test_non_finite_coordinates_rejected(cuda):
    """A NaN or Inf coordinate stops the build at the neighbour search with INCH_ERR_NAN."""
    from inching_amd._lib import InchingError
    from inching_amd.hessian import build_hessian
    print("case: non-finite coordinates", flush=True)
    for bad in (np.nan, np.inf):
        X = small_globule(300, seed=64)
        X[17, 1] = bad
        with pytest.raises(InchingError) as e:
            build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
        assert e.value.status == -5


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained weights not present")
def test_corrupted_network_falls_back_to_cold_start(cuda):
    """A network whose weights hold NaN predicts a non-finite block: both warm pipelines use
    the cold start instead (fallback=True) and return the dense-eigh eigenvalues; a finite
    but untrained network is still only a start and also returns them."""
    from inching_amd import warmstart as W
    from inching_amd.hessian import build_hessian, rigid_body_basis
    from inching_amd.pipeline import lowest_modes_jdm_warm, lowest_modes_warm
    print("case: corrupted warm block", flush=True)
    X = small_globule(1000, seed=65)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=())
    Q = rigid_body_basis(A.xyz)
    w = np.linalg.eigvalsh(OH.hessian_dense(X[A.order], 8.0, 1.0))
    bad = W.WarmStartNet.load()
    with torch.no_grad():
        bad.params[0].fill_(float("nan"))
    untrained = W.WarmStartNet(channels=bad.C, seed=5)
    k = 10
    for net, expect_fallback in ((bad, True), (untrained, None)):
        for solve in (lowest_modes_warm, lowest_modes_jdm_warm):
            r = solve(A, k, net, Q=Q, precision="mf64", tol_rel=1e-9)
            if expect_fallback is not None:
                assert r.fallback is expect_fallback
            assert r.converged and np.max(np.abs(r.eigval - w[6:6 + k])) <= 1e-9
