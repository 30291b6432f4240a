"""Drop-in for ``InchingLiteInteger/Burn/HermitianLanczos/T2.py``: spectrum bound.

``A_Adiag_EstimateSpectrumBound(A, A_diag, User_HalfMemMode=True)`` (reference :30-101):
(min, max) Ritz values of a 300-step Lanczos run without re-orthogonalisation, returned as
floats that also answer ``.get()`` like the reference's 0-d device arrays.
"""
from inching_amd import chebyshev as CH

from ..._common import DevScalar, require_block_matrix


def A_Adiag_EstimateSpectrumBound(A, A_diag, User_HalfMemMode=True, seed=0):
    A = require_block_matrix(A)
    lo, hi = CH.spectrum_bound(A, precision="f64" if A.vals64 is not None else "f32", seed=seed)
    return DevScalar(lo), DevScalar(hi)
