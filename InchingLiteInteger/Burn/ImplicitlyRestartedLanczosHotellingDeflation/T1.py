"""Drop-in for ``InchingLiteInteger/Burn/ImplicitlyRestartedLanczosHotellingDeflation/T1.py``.

``S_HeigvalIRLMHD_HeigvecIRLMHD(A, k=32, maxiter=15000, tol=1e-8, ...)`` (reference
:308-477) returns ``(eigval[k], eigvec[n, k])`` as GPU tensors. The reference's driver
builds its matvec with a single argument (``:368,370``) and raises TypeError as shipped;
this version takes the diagonal from the matrix (``A_diag`` may also be passed).
"""
import torch

from inching_amd.irlm import IRLM
from inching_amd.operator import Operator

from ..._common import dense_Q, require_block_matrix


def S_HeigvalIRLMHD_HeigvecIRLMHD(A, k=32, maxiter=15000, tol=1e-8, User_HalfMemMode=True,
                                  User_Q_HotellingDeflation=None, User_HotellingShift=10.0, A_diag=None,
                                  precision="f64", seed=0):
    A = require_block_matrix(A)
    n = A.shape[0]
    assert k < n, "ABORTED. k must be smaller than n"
    assert len(A.shape) == 2, "ABORTED. It is a tensor not rank 2!"
    assert A.shape[0] == A.shape[1], "ABORTED. square"
    if User_Q_HotellingDeflation is None:
        print("WARNING. Hotelling deflation not in use")
    Q = dense_Q(User_Q_HotellingDeflation, A.device)
    op = Operator(A, precision, Q=Q, shift=User_HotellingShift)
    r = IRLM(op, k).run(maxiter=maxiter, tol=tol, seed=seed)
    return torch.as_tensor(r.eigval, device=A.device), r.eigvec.T
