"""Drop-in for ``InchingLiteInteger/Burn/PolynomialFilters/T2.py``: filtered matvec, KPM
count and the interval bisection.

* ``OOC2_A_Adiag_ChebyshevAv(A, A_diag, User_PolynomialParams, User_HalfMemMode)``
  (reference :16-92) with ``.ChebyshevAv(A, v, User_ReturnRho=False)``: the degree-m filter,
  one fused SpMV + three-term-recurrence kernel per degree; with User_ReturnRho the KPM
  density term is accumulated from the per-degree dots v . T_k v.
* ``A_Adiag_PolynomialParams_KpmEstimate`` (:96-138): int(mean rho * n) + 1 over the probes.
* ``A_Adiag_OptimizePolynomialParamsOnMemory`` (:142-257): bisection of the right end of
  the wanted interval; the probe moments are computed once and reused by every step.
"""
import numpy as np
import torch

from inching_amd import chebyshev as CH
from inching_amd import ops
from inching_amd.operator import Operator

from ..._common import require_block_matrix


class OOC2_A_Adiag_ChebyshevAv:
    def __init__(self, A, A_diag, User_PolynomialParams=None, User_HalfMemMode=True):
        assert User_PolynomialParams is not None, "ABORTED. You must initiate with a found Polynomial Param"
        self.User_PolynomialParams = User_PolynomialParams
        self.A = require_block_matrix(A)
        self.N = A.shape[0]
        self.precision = "f64" if A.vals64 is not None else "f32"

    def ChebyshevAv(self, A, v, User_ReturnRho=False):
        pol = self.User_PolynomialParams
        vv = v.reshape(-1).to(torch.float64).contiguous()
        op = Operator(self.A, self.precision, pol=pol)
        y = torch.empty_like(vv)
        if not User_ReturnRho:
            op.filtered(vv, y)
            return y
        # recurrence with the per-degree dot products for the KPM term
        m, dd, cc = int(pol.AdjustedDegree), float(pol.dd), float(pol.cc)
        mu = pol.AdjustedCoefficients
        jac = pol.AdjustedDampingKernels
        vk, vkm1, vkp1 = vv.clone(), torch.zeros_like(vv), torch.empty_like(vv)
        acc = vv * float(mu[0])
        ws = ops.Workspace(1, vv.device)
        dot = torch.empty(1, dtype=torch.float64, device=vv.device)
        rho = pol.theta_a - pol.theta_b
        for k in range(1, m + 1):
            if k == 1:
                op.spmv(vk, vkp1, s1=1.0 / dd, c0=cc, acc=acc, mu=float(mu[1]))
            else:
                op.spmv(vk, vkp1, s1=2.0 / dd, c0=cc, z=vkm1, c2=1.0, acc=acc, mu=float(mu[k]))
            vkm1, vk, vkp1 = vk, vkp1, vkm1
            ops.multidot(vk.view(1, -1), 1, vv, dot, ws)
            rho -= 2 * jac[k] * dot.item() * (np.sin(k * pol.theta_b) - np.sin(k * pol.theta_a)) / k
        rho /= np.pi
        return acc, rho


def A_Adiag_PolynomialParams_KpmEstimate(A, A_diag, User_HalfMemMode=True, User_PolynomialParams=None,
                                         User_NumberKpmTrials=1000, seed=0):
    assert User_PolynomialParams is not None, "ABORTED. You must initiate with a found Polynomial Param"
    A = require_block_matrix(A)
    pol = User_PolynomialParams
    lo = pol.cc - pol.dd
    hi = pol.cc + pol.dd
    kpm = CH.KPMMoments(A, (lo, hi), trials=User_NumberKpmTrials, seed=seed,
                        precision="f64" if A.vals64 is not None else "f32")
    return CH.kpm_count_from_moments(pol, kpm.moments(int(pol.AdjustedDegree)), A.shape[0])


def A_Adiag_OptimizePolynomialParamsOnMemory(A, A_diag, User_MaximumDegree=1000, User_MinimumDegree=10,
                                             User_DampingKernel="Jackson", User_ExtremalIntervalDefinition=1e-10,
                                             User_WantedInterval=(36.8, 37.2), User_SpectrumBound=(1.0, 759.0),
                                             User_DesignatedStart=None, User_WantedNumberEigenvalue=12,
                                             User_AffordableMemoryMargin=5, User_HalfMemMode=True,
                                             User_NumberKpmTrials=1000, User_ConvergenceRatio=0.1, seed=0):
    A = require_block_matrix(A)
    pol, cnt, interval, _ = CH.optimize_on_memory(
        A, (float(User_SpectrumBound[0]), float(User_SpectrumBound[1])), float(User_WantedInterval[0]),
        int(User_WantedNumberEigenvalue), designated_start=User_DesignatedStart, max_degree=User_MaximumDegree,
        min_degree=User_MinimumDegree, kernel=User_DampingKernel, ext=User_ExtremalIntervalDefinition,
        margin=User_AffordableMemoryMargin, trials=User_NumberKpmTrials, ratio=User_ConvergenceRatio, seed=seed,
        precision="f64" if A.vals64 is not None else "f32")
    return pol, cnt, interval
