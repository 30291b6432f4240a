"""Drop-in for ``InchingLiteInteger/Burn/ChebyshevDavidsonSubspaceIteration/T1.py``.

``S_HeigvalCDSIHD_HeigvecCDSIHD`` (reference :229-623) keeps its arguments, asserts and
return convention ``(eigval[k] ascending, eigvec[n, k])`` (GPU tensors). The spectrum bound
is required and Hotelling deflation is rejected, as in the reference; the lowest k
eigenpairs (rigid-body modes included) are returned. A run that does not converge within
``maxiter`` prints the reference's message and returns what it has.
"""
import torch

from inching_amd.cdsi import CDSI

from ..._common import require_block_matrix


def _float(v):
    return float(v.get()) if hasattr(v, "get") else float(v)


def S_HeigvalCDSIHD_HeigvecCDSIHD(A, A_diag, k=16, tol=1e-10, maxiter=1000000, User_WorkspaceSizeFactor=2,
                                  User_ChebyshevDegree=80, User_HalfMemMode=True, User_SpectrumBound=None,
                                  User_Q_HotellingDeflation=None, User_HotellingShift=10.0, precision="mf64",
                                  seed=0):
    assert User_SpectrumBound is not None, "ABORTED. The spectrum bound has to be initiated."
    assert User_WorkspaceSizeFactor > 1, "ABORTED. We require User_WorkspaceSizeFactor > 1 "
    assert User_Q_HotellingDeflation is None, (
        "ABORTED. We do not support Hotelling at the moment (Actually we do not need it if we have a mid-pass "
        "filter.) ")
    print("WARNING. Hotelling deflation not in use")
    A = require_block_matrix(A)
    bound = (_float(User_SpectrumBound[0]), _float(User_SpectrumBound[-1]))
    r = CDSI(A, k, factor=User_WorkspaceSizeFactor, degree=User_ChebyshevDegree, precision=precision).run(
        bound, tol=tol, maxiter=maxiter, seed=seed)
    if r.converged < k:
        print("ABORTED. It did not actually converged! We went through %s coarse iter and collected %s converged"
              % (r.iters, r.converged))
    return torch.as_tensor(r.eigval, device=A.device), r.eigvec.T
