"""FP64 Rayleigh-Ritz polish and residual check of a block of approximate eigenvectors.

The reference checks each returned pair with ||A v - lam v|| on its float64 operator
(Command8A/commandBenchmark_InchingTRLMHD_Linux_Pdb_0064_A100.py:470-476). Here the float64
operator is the matrix-free ANM product recomputed from float64 coordinates and the
neighbour pattern (no float64 copy of the Hessian values), so a float32 solve can be
polished and verified in float64: the block is cleaned of rigid-body components,
orthonormalised (two Cholesky-QR passes), and a Rayleigh-Ritz step in float64 gives
eigenvalues with second-order error and the residual norms.
"""
import numpy as np
import torch

from . import ops


def _gram(X, Y, ws=None):
    """G[a, b] = X[a] . Y[b] for row blocks X [k1, n], Y [k2, n] (float64 MFMA kernel) -> numpy."""
    return ops.gram(X, Y, ws=ws).cpu().numpy()


def _combine(X, W):
    """out[q] = sum_r W[r, q] X[r] (W numpy [m, k])."""
    out = torch.empty((W.shape[1], X.shape[1]), dtype=torch.float64, device=X.device)
    ops.rr_combine(X, W.shape[0], torch.as_tensor(np.ascontiguousarray(W), device=X.device), out)
    return out


def orthonormalise(X, Q=None, passes=2):
    """Rows of X made orthonormal (and orthogonal to the rows of Q when given)."""
    k = X.shape[0]
    m = 0 if Q is None else Q.shape[0]
    ws = ops.GramWorkspace(max(k, m, 1), max(k, m, 1), X.device)
    if Q is not None:
        for _ in range(2):
            C = ops.gram(X, Q, ws=ws)                     # [k, m]
            P = torch.empty_like(X)
            ops.rr_combine(Q, m, C.T.contiguous(), P)     # P[q] = sum_r C[q, r] Q[r]
            ops.axpby(-1.0, P.view(-1), 1.0, X.view(-1))
    for _ in range(passes):
        G = _gram(X, X, ws)
        G = 0.5 * (G + G.T)
        R = np.linalg.cholesky(G).T          # G = R^T R
        X = _combine(X, np.linalg.inv(R))
    return X


def apply_fp64(A, X, operator="matfree"):
    """Y = A X for a row block X [k, n] in float64."""
    Xc = X.T.contiguous()
    if operator == "matfree":
        Yc = A.matfree(Xc)
    else:
        Yc = A.matmat(Xc, precision="f64")
    return Yc.T.contiguous()


def residual_norms(X, Y, theta):
    """||Y[j] - theta_j X[j]||_2 for each row j (one fused kernel)."""
    th = torch.as_tensor(np.asarray(theta, dtype=np.float64), device=X.device)
    return ops.residual_norms(X, Y, th).cpu().numpy()


def polish(A, Vrows, Q=None, operator="matfree", lam_max=None):
    """Float64 Rayleigh-Ritz of the row block Vrows [k, n]. Returns (eigenvalues ascending,
    eigenvectors as rows [k, n], info dict with absolute and relative residuals)."""
    X = orthonormalise(Vrows.to(torch.float64).contiguous().clone(), Q=Q)
    Y = apply_fp64(A, X, operator)
    Hm = _gram(X, Y)
    Hm = 0.5 * (Hm + Hm.T)
    theta, C = np.linalg.eigh(Hm)
    X = _combine(X, C)
    Y = _combine(Y, C)
    res = residual_norms(X, Y, theta)
    info = {"res": res}
    if lam_max is not None:
        info["rel_res"] = res / float(lam_max)
    return theta, X, info
