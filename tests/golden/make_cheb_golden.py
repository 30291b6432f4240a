"""Write golden Chebyshev-filter design vectors by running the reference's own design code.

The reference module ``InchingLiteInteger/Burn/PolynomialFilters/T0.py`` imports numpy
only, so it runs in this container. Nothing else of the reference is imported. The
module is loaded from its file path (no package import) and called on a fixed set of
cases; the inputs and outputs are stored as plain arrays in ``cheb_design.npz``.

Run from the repository root:  python tests/golden/make_cheb_golden.py
"""
import importlib.util
import json
import os
import sys

import numpy as np

REF = "/root/reference/InchingLiteInteger/Burn/PolynomialFilters/T0.py"
HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [
    # (MaxDeg, MinDeg, kernel, ext, wanted, bound, ratio)
    (5120, 5, "Jackson", 1e-10, (1.0, 4.0), (1.0, 40.0), 0.7),
    (5120, 5, "Jackson", 1e-10, (1.0, 1.3), (1.0, 40.0), 0.7),
    (1000, 10, "Jackson", 1e-9, (1.0, 2.5), (1.0, 25.0), 0.1),
    (1000, 10, "LanczosSigma", 1e-9, (1.0, 2.5), (1.0, 25.0), 0.1),
    (1000, 10, "Jackson", 1e-9, (30.0, 40.0), (1.0, 40.0), 0.1),
    (1000, 10, "Jackson", 1e-9, (5.0, 7.0), (1.0, 40.0), 0.1),
    (1000, 10, "Jackson", 1e-9, (5.0, 7.0), (1.0, 40.0), 0.7),
    (2000, 5, "LanczosSigma", 1e-9, (12.0, 14.5), (1.0, 60.0), 0.3),
    (1000, 10, "None", 1e-9, (10.0, 12.0), (1.0, 40.0), 0.2),
]


def main():
    spec = importlib.util.spec_from_file_location("ref_t0", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {}
    meta = []
    for ci, (maxd, mind, ker, ext, wanted, bound, ratio) in enumerate(CASES):
        pol = mod.OOC0_OptimizePolynomialParamsOnDegree(
            User_MaximumDegree=maxd, User_MinimumDegree=mind, User_DampingKernel=ker,
            User_ExtremalIntervalDefinition=ext, User_WantedInterval=wanted,
            User_SpectrumBound=bound, User_ConvergenceRatio=ratio)
        out["mu_%d" % ci] = np.asarray(pol.AdjustedCoefficients, dtype=np.float64)
        out["jac_%d" % ci] = np.asarray(pol.AdjustedDampingKernels, dtype=np.float64)
        meta.append(dict(case=[maxd, mind, ker, ext, list(wanted), list(bound), ratio],
                         degree=int(pol.AdjustedDegree), gam=float(pol.gam),
                         theta_a=float(pol.theta_a), theta_b=float(pol.theta_b),
                         cc=float(pol.cc), dd=float(pol.dd)))
        # the damping kernel on its own for a few degrees
    for m in (5, 17, 100):
        for ker in ("Jackson", "LanczosSigma", "None"):
            out["damp_%s_%d" % (ker, m)] = mod.OOC1_DegreeM_DampingKernel(m, User_MaximumDegree=200, User_DampingKernel=ker)
    out["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "cheb_design.npz"), **out)
    print("wrote", len(meta), "cases")


if __name__ == "__main__":
    sys.exit(main())
