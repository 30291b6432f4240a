"""Drop-in for ``InchingLiteInteger/Burn/ThickRestartLanczosHotellingDeflation/T1.py``.

``S_HeigvalTRLMHD_HeigvecTRLMHD`` (reference :60-310) keeps its arguments and returns
``(eigval[k] ascending, eigvec[n, k])`` as GPU tensors; eigenvalues include PlusI. The
solver is ``inching_amd.trlm.TRLM`` on the block SpMV kernels: full CGS2, thick restart
with the arrowhead Rayleigh-Ritz, Hotelling term for ``User_Q_HotellingDeflation``, and with
``User_PolynomialParams`` the Chebyshev-filtered operator (largest k of p(A) selected, then
Rayleigh quotients with A, reference :277-291).

Extra keyword arguments (not in the reference): ``precision`` ('f64' default, 'f32' or
'mf64' matrix-free), ``v0`` (start vector) and ``seed``.
"""
import numpy as np
import torch

from inching_amd.operator import Operator
from inching_amd.trlm import TRLM

from ..._common import dense_Q, require_block_matrix


def S_HeigvalTRLMHD_HeigvecTRLMHD(A, A_diag, k=32, maxiter=10000, tol=0, User_HalfMemMode=True,
                                  User_WorkspaceSizeFactor=2, User_Q_HotellingDeflation=None,
                                  User_HotellingShift=10.0, User_PolynomialParams=None, precision="f64", v0=None,
                                  seed=0):
    A = require_block_matrix(A)
    N = A.shape[0]
    assert k < N, "ABORTED. k must be smaller than n"
    assert len(A.shape) == 2, "ABORTED. It is a tensor not rank 2!"
    assert A.shape[0] == A.shape[1], "ABORTED. square"
    if User_Q_HotellingDeflation is None:
        print("WARNING. Hotelling deflation not in use")
    Q = dense_Q(User_Q_HotellingDeflation, A.device)
    op = Operator(A, precision, pol=User_PolynomialParams, Q=Q, shift=User_HotellingShift)
    solver = TRLM(op, k, factor=User_WorkspaceSizeFactor)
    r = solver.run(maxiter=maxiter, tol=tol, v0=v0, seed=seed, filtered=User_PolynomialParams is not None)
    S = np.asarray(r.eigval)
    V = r.eigvec
    if User_PolynomialParams is not None:
        plain = Operator(A, precision)
        u = torch.empty(N, dtype=torch.float64, device=A.device)
        S = np.array([float(torch.dot(V[i], plain.plain(V[i], u))) for i in range(V.shape[0])])
    idx = np.argsort(S)
    idx_t = torch.as_tensor(idx, device=A.device)
    return torch.as_tensor(S[idx], device=A.device), V[idx_t].T
