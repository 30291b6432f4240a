"""ANM Hessian, rigid-body basis and the half-storage matvec (oracle).

Hessian: ``InchingLiteInteger/Fuel/T1.py:261-456`` (``ReturnCupyHLowerTriangleInt32``):
for neighbours j != i with 0 < r^2 <= rc^2 and d = x_j - x_i (:276),
H_ij = -d d^T / r^2 (:281, :291-307), H_ii = -sum_{j != i} H_ij (:410-414), and
A = H + PlusI * I (:416-421). The reference returns the scalar lower triangle
(including the diagonal) of A as CSR (:444, :655-660) plus the full scalar diagonal
``A_diag`` (:445-449).

Rigid-body basis: driver code ``Command8A/commandBenchmark_InchingTRLMHD_Linux_Pdb_0064_A100.py:314-372``.

Matvec: ``InchingLiteInteger/Burn/Krylov/T3.py:54-165`` (``u = L v + L^T v - A_diag * v``).
Hotelling/EED: ``InchingLiteInteger/Burn/Orthogonalization/T3.py:62-164``
(``res += shift * Q^T (Q x)``).
"""
import numpy as np
import scipy.sparse as sp

from .geometry import neighbour_csr


def block_hessian(X, rc, plus_i=1.0, nb=None):
    """Full-storage 3x3 block rows: (rowptr int64[N+1], col int32[nnzb], vals float64[nnzb,3,3]).

    Columns ascending within each row; the diagonal block carries PlusI.
    """
    X = np.ascontiguousarray(X, dtype=np.float64)
    n = X.shape[0]
    if nb is None:
        nb = neighbour_csr(X, rc)
    indptr, indices = nb
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    rows = np.repeat(np.arange(n), np.diff(indptr))
    d = X[indices] - X[rows]
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    rc2 = float(rc) * float(rc)
    keep = (r2 > 0) & (r2 <= rc2)
    coeff = np.zeros_like(r2)
    coeff[keep] = -1.0 / r2[keep]
    vals = d[:, :, None] * d[:, None, :] * coeff[:, None, None]
    # diagonal block = -(row sum of off-diagonal blocks) + PlusI
    diag = np.zeros((n, 3, 3))
    np.add.at(diag, rows, -vals)
    diag += plus_i * np.eye(3)[None]
    selfpos = indices == rows
    vals[selfpos] = diag[rows[selfpos]]
    return indptr, indices.astype(np.int32), vals


def block_to_dense(rowptr, col, vals, n_atoms):
    A = np.zeros((3 * n_atoms, 3 * n_atoms))
    rows = np.repeat(np.arange(n_atoms), np.diff(rowptr))
    for b in range(col.shape[0]):
        i, j = rows[b], col[b]
        A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = vals[b]
    return A


def block_to_scipy(rowptr, col, vals, n_atoms):
    return sp.bsr_matrix((vals, np.asarray(col, dtype=np.int64), rowptr), shape=(3 * n_atoms, 3 * n_atoms)).tocsr()


def hessian_dense(X, rc, plus_i=1.0):
    rp, c, v = block_hessian(X, rc, plus_i)
    return block_to_dense(rp, c, v, X.shape[0])


def lower_csr_and_diag(X, rc, plus_i=1.0, nb=None):
    """The reference's return pair: scalar CSR of tril(A) (exact zeros dropped) and A_diag."""
    rp, c, v = block_hessian(X, rc, plus_i, nb=nb)
    A = block_to_scipy(rp, c, v, X.shape[0])
    L = sp.tril(A, k=0, format="csr")
    L.eliminate_zeros()
    L.sort_indices()
    return L, A.diagonal().copy()


def halfmem_matvec(L, A_diag, v):
    """Krylov/T3.py:131-160: u = L v; u += L^T v; u -= A_diag * v."""
    u = L @ v
    u += L.T @ v
    u -= A_diag * v
    return u


def rigid_body_basis(X):
    """6 x 3N rigid-body basis exactly as the driver builds it (two double-MGS sweeps)."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    Q = np.zeros((6, 3 * n))
    for a in range(3):
        q = np.zeros((n, 3))
        q[:, a] = 1.0 / np.sqrt(n)
        Q[a] = q.ravel()
    Rx = np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0]], dtype=np.float64).T
    Ry = np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0]], dtype=np.float64).T
    Rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 0]], dtype=np.float64).T
    for row, R in zip((3, 4, 5), (Rx, Ry, Rz)):
        g = (X @ R).ravel()
        Q[row] = g / np.linalg.norm(g)
    for _ in range(2):
        V = Q.T.copy()
        for ix in range(1, 6):
            for _rep in range(2):
                V[:, ix] -= V[:, :ix] @ (V[:, :ix].T @ V[:, ix])
                V[:, ix] /= np.sqrt(V[:, ix] @ V[:, ix])
        Q = V.T.copy()
    return Q


def hotelling(Q, x, res, shift):
    """Orthogonalization/T3.py:62-164: res += shift * Q^T (Q x) (in place, returns res)."""
    res += shift * (Q.T @ (Q @ x))
    return res


def relative_residual(A_apply, lam, V, lam_max):
    """||A v - lam v||_2 / (lam_max ||v||_2) per column (the build's rel_res definition)."""
    R = A_apply(V) - V * lam[None, :]
    return np.linalg.norm(R, axis=0) / (lam_max * np.linalg.norm(V, axis=0))


def block_hessian_pbc(X, rc, plus_i, box):
    """Periodic x/y images (``Fuel/Coordinate/T1.py:67-161``, ``Fuel/T1.py:98-128,317-396``),
    brute force over all pairs and the nine images s = (sx, sy) in {-1,0,1}^2:
    d_s = (x_j + s * box) - x_i in float64, j a neighbour of i when some image has
    |d_s|^2 <= rc^2 (self included), H_ij = sum over those images of -d_s d_s^T / |d_s|^2
    (|d_s|^2 > 0), H_ii = -sum_{j != i} H_ij + plus_i I. Returns (rowptr, col, vals [nnzb,3,3],
    image masks, bit (sx+1)*3 + (sy+1))."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    rc2 = float(rc) * float(rc)
    shifts = [(sx, sy) for sx in (-1, 0, 1) for sy in (-1, 0, 1)]
    rows, cols, blocks, masks = [], [], [], []
    rowptr = np.zeros(n + 1, dtype=np.int64)
    for i in range(n):
        dz = X[:, 2] - X[i, 2]
        m = np.zeros(n, dtype=np.int64)
        dd = {}
        for bit, (sx, sy) in enumerate(shifts):
            dx = (X[:, 0] + sx * box[0]) - X[i, 0]
            dy = (X[:, 1] + sy * box[1]) - X[i, 1]
            d2 = (dx * dx + dy * dy) + dz * dz
            hit = d2 <= rc2
            m[hit] |= 1 << bit
            dd[bit] = (dx, dy, d2)
        js = np.nonzero(m)[0]
        diag = np.zeros((3, 3))
        self_pos = None
        for j in js:
            if j == i:
                self_pos = len(blocks)
                blocks.append(None)
            else:
                h = np.zeros((3, 3))
                for bit in range(9):
                    if (m[j] >> bit) & 1:
                        dx, dy, d2 = dd[bit]
                        if d2[j] > 0:
                            d = np.array([dx[j], dy[j], dz[j]])
                            h -= np.outer(d, d) / d2[j]
                blocks.append(h)
                diag -= h
            cols.append(j)
            masks.append(int(m[j]))
        blocks[self_pos] = diag + plus_i * np.eye(3)
        rowptr[i + 1] = len(cols)
    return rowptr, np.asarray(cols, dtype=np.int32), np.asarray(blocks), np.asarray(masks, dtype=np.int64)
