"""Drop-in for ``InchingLiteInteger/Burn/PolynomialFilters/T0.py``: Chebyshev filter design.

Same names, arguments and float64 results as the reference module (PolParams :4-20,
OOC1_DegreeM_DampingKernel :25-55, the extremal and interior designers :64-169 / :248-303,
rootchb :182-226, ChebIv :229-245, OOC0_OptimizePolynomialParamsOnDegree :313-395); the
implementation lives in ``inching_amd.chebyshev``.
"""
import numpy

from inching_amd.chebyshev import (ChebIv, OOC0_OptimizePolynomialParamsOnDegree, OOC1_DegreeM_DampingKernel,  # noqa: F401
                                   PolParams, _balance, _extremal, _interior)


def OOC3_PolynomialParams_aa_bb_ExtremalPolynomialParams(pol, a, b):
    return _extremal(pol, a, b)


def OOC3_PolynomialParams_aa_bb_InteriorPolynomialParams(pol, aa, bb):
    return _interior(pol, aa, bb)


def dif_eval(v, thc, jac):
    return numpy.sum(v * numpy.cos(numpy.arange(jac.shape[0]) * thc, dtype=numpy.float64) * jac)


def rootchb(m, v, jac, tha, thb):
    return _balance(m, v, jac, tha, thb)
