"""Neighbour sets (oracle).

Reference: the k-d tree ball queries of ``InchingLiteInteger/Fuel/Coordinate/T1.py``
(:219, :239, :276-277) and the Hessian cutoff mask of ``InchingLiteInteger/Fuel/T1.py``
(:291-299): atom j is a neighbour of atom i iff |x_j - x_i|^2 <= rc^2. Self is
included (degree counts it, ``Fuel/Coordinate/T1.py:236-245``).

Pinned rule used here (the reference leaves the floating-point evaluation
to scipy/CuPy): d^2 = (dx*dx + dy*dy) + dz*dz in float64 with dx = x_j - x_i, no
fused multiply-add, compared with rc*rc (rc*rc evaluated in float64).
"""
import numpy as np


def _exact_mask(X, i, cand, rc2):
    d = X[cand] - X[i]
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return cand[d2 <= rc2]


def neighbour_csr(X, rc):
    """Return (indptr int64 [N+1], indices int32) with ascending columns, self included."""
    from scipy.spatial import cKDTree
    X = np.ascontiguousarray(X, dtype=np.float64)
    n = X.shape[0]
    rc2 = float(rc) * float(rc)
    tree = cKDTree(X)
    cands = tree.query_ball_point(X, float(rc) * (1.0 + 1e-9) + 1e-12, return_sorted=True)
    rows = []
    for i in range(n):
        c = np.asarray(cands[i], dtype=np.int64)
        rows.append(np.sort(_exact_mask(X, i, c, rc2)))
    indptr = np.zeros(n + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    indices = np.concatenate(rows).astype(np.int32) if n else np.zeros(0, np.int32)
    return indptr, indices


def neighbour_csr_bruteforce(X, rc):
    """All-pairs O(N^2) version of :func:`neighbour_csr` (small N only)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n = X.shape[0]
    rc2 = float(rc) * float(rc)
    rows = []
    allj = np.arange(n)
    for i in range(n):
        rows.append(_exact_mask(X, i, allj, rc2))
    indptr = np.zeros(n + 1, dtype=np.int64)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    indices = np.concatenate(rows).astype(np.int32) if n else np.zeros(0, np.int32)
    return indptr, indices


def degrees(indptr):
    return np.diff(indptr).astype(np.int32)
