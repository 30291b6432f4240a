"""Drop-in for ``InchingLiteInteger/Burn/JacobiDavidsonHotellingDeflation/T1.py``.

``S_HeigvalJDMHD_HeigvecJDMHD`` (reference :382-757) keeps its arguments and returns
``(eigval[k] ascending, eigvec[n, k])`` as GPU tensors. As in the reference a polynomial
filter is rejected (the correction equation targets the smallest eigenvalues), and a run
that does not converge within ``maxiter`` prints the reference's message and returns what
it has.
"""
import torch

from inching_amd.jdm import JDM
from inching_amd.operator import Operator

from ..._common import dense_Q, require_block_matrix


def S_HeigvalJDMHD_HeigvecJDMHD(A, A_diag, k=16, tol=1e-10, maxiter=1000, User_CorrectionSolverMaxiter=20,
                                User_WorkspaceSizeFactor=2, User_HalfMemMode=True,
                                User_IntermediateConvergenceTol=1e-3, User_GapEstimate=0.1,
                                User_FactoringToleranceOnCorrection=1e-4, User_Q_HotellingDeflation=None,
                                User_HotellingShift=10.0, User_PolynomialParams=None, precision="f64", seed=0):
    assert User_PolynomialParams is None, (
        "ABORTED. WARNING. The J in JDM is targeting the smallest eigval due to implicit inversion which "
        "counteract the goal of Cheb to map to the largest eigval. You are advised to Use CDSI instead.")
    assert User_WorkspaceSizeFactor > 1, "ABORTED. User_WorkspaceSizeFactor: int > 1"
    A = require_block_matrix(A)
    if User_Q_HotellingDeflation is None:
        print("WARNING. Hotelling deflation not in use")
    Q = dense_Q(User_Q_HotellingDeflation, A.device)
    op = Operator(A, precision, Q=Q, shift=User_HotellingShift)
    r = JDM(op, k, factor=User_WorkspaceSizeFactor).run(
        tol=tol, maxiter=maxiter, corr_maxiter=User_CorrectionSolverMaxiter,
        inter_tol=User_IntermediateConvergenceTol, gap=User_GapEstimate,
        corr_factor=User_FactoringToleranceOnCorrection, seed=seed)
    if r.converged < k:
        print("ABORTED. It did not actually converged! We went through %s coarse iter and collected %s converged"
              % (r.iters, r.converged))
    return torch.as_tensor(r.eigval, device=A.device), r.eigvec.T
