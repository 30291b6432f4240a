"""Chebyshev filter design, filtered matvec and KPM count (oracle).

Design restates ``InchingLiteInteger/Burn/PolynomialFilters/T0.py`` step for step (the
golden vectors in ``tests/golden/cheb_design.npz`` were written by the reference's own
T0 module, so the arithmetic order below is kept as the reference evaluates it):

* ``PolParams``                       T0.py:4-20
* ``damping_kernel``                  T0.py:25-55 (OOC1_DegreeM_DampingKernel)
* ``extremal_params``                 T0.py:64-169 (OOC3_..._ExtremalPolynomialParams)
* ``_dif_eval`` / ``rootchb``         T0.py:172-226
* ``cheb_iv``                         T0.py:229-245 (ChebIv)
* ``interior_params``                 T0.py:248-303 (OOC3_..._InteriorPolynomialParams)
* ``optimize_on_degree``              T0.py:313-395 (OOC0_OptimizePolynomialParamsOnDegree)

Filtered matvec / KPM / bisection restate ``Burn/PolynomialFilters/T2.py``:
``cheb_av`` :28-92, ``kpm_estimate`` :96-138, ``optimize_on_memory`` :142-257.
"""
import numpy as np


class PolParams:
    def __init__(self, User_MaximumDegree=10000, User_MinimumDegree=2, User_DampingKernel="Jackson",
                 mu=None, deg=0, User_ExtremalIntervalDefinition=1e-9, User_ConvergenceRatio=0.1):
        self.User_MaximumDegree = User_MaximumDegree
        self.User_MinimumDegree = User_MinimumDegree
        self.User_DampingKernel = User_DampingKernel
        self.User_ExtremalIntervalDefinition = User_ExtremalIntervalDefinition
        self.User_ConvergenceRatio = User_ConvergenceRatio
        self.AdjustedCoefficients = mu
        self.AdjustedDampingKernels = None
        self.AdjustedDegree = deg


def damping_kernel(m, max_degree=1000, kind="Jackson"):
    jac = np.zeros(max_degree + 1, dtype=np.float64)
    k = np.arange(1, m + 1, dtype=np.float64)
    if kind == "Jackson":
        th = np.pi / (m + 2)
        a1 = 1 / (m + 2)
        a2 = np.sin(th, dtype=np.float64)
        jac[1:m + 1] = a1 * np.sin((k + 1) * th, dtype=np.float64) / a2 + (1 - (k + 1) * a1) * np.cos(k * th, dtype=np.float64)
    elif kind == "LanczosSigma":
        th = np.pi / (m + 1)
        jac[1:m + 1] = np.sin(k * th, dtype=np.float64) / (k * th)
    else:
        jac[1:] = 1.0
    jac[0] = 0.5
    return jac


def extremal_params(pol, a, b):
    maxdeg = pol.User_MaximumDegree
    mindeg = pol.User_MinimumDegree
    gam = np.sign(b - (1 - pol.User_ExtremalIntervalDefinition))
    x = a if gam > 0 else b
    a, b = tuple(sorted([x - gam * 0.1 * np.sqrt((1 - gam * x) / 2, dtype=np.float64), -1 * gam]))
    # degree: run the scalar recurrence until the value at x drops under 0.1
    c = (b + a) * 0.5
    e = (b - a) * 0.5
    sigma = e / (gam - c)
    sigma1 = sigma
    g0 = 1.0
    g1 = (x - c) * (sigma1 / e)
    m = 1
    for k in range(2, maxdeg + 1):
        sigma_new = 1.0 / (2.0 / sigma1 - sigma)
        s1 = sigma_new / e
        gnew = 2 * (x - c) * s1 * g1 - (sigma * sigma_new) * g0
        g0 = np.copy(g1)
        g1 = np.copy(gnew)
        sigma = sigma_new
        m += 1
        if (g1 < 0.1) & (k >= mindeg):
            break
    bar = np.copy(g1)
    # coefficients: the same recurrence carried on Chebyshev coefficient vectors
    sigma = e / (gam - c)
    sigma1 = sigma
    tau = 2 / sigma1
    G0 = np.zeros(maxdeg + 1, dtype=np.float64)
    G1 = np.zeros(maxdeg + 1, dtype=np.float64)
    GN = np.zeros(maxdeg + 1, dtype=np.float64)
    G0[0] = 1.0
    G1[0] = -c * (sigma1 / e)
    G1[1] = 1 * (sigma1 / e)
    for k in range(2, m + 1):
        sigma_new = 1.0 / (tau - sigma)
        s1 = sigma_new / e
        GN[:k] = G1[1:k + 1]          # x T_j -> T_{j-1} / 2 part (times 2)
        GN[1] += G1[0]                # x T_0 = T_1 (times 2)
        GN[1:k + 1] += G1[:k]         # x T_j -> T_{j+1} / 2 part (times 2)
        GN *= s1
        GN[:k + 1] -= (2 * s1 * c) * G1[:k + 1]
        GN[:k + 1] -= (sigma * sigma_new) * G0[:k + 1]
        G0[:k + 1] = np.copy(G1[:k + 1])
        G1[:k + 1] = np.copy(GN[:k + 1])
        sigma = sigma_new
    pol.AdjustedCoefficients = G1[:m + 1] * pol.AdjustedDampingKernels[:m + 1]
    pol.AdjustedDegree = m
    pol.bar = bar
    pol.gam = gam
    return pol


def _dif_eval(v, thc, jac):
    return np.sum(v * np.cos(np.arange(jac.shape[0]) * thc, dtype=np.float64) * jac)


def rootchb(m, v, jac, tha, thb):
    """Newton steps safeguarded by bisection for the balancing angle."""
    tol_bal = abs(tha - thb) * 1e-6
    thc = (tha + thb) / 2
    for _ in range(759):
        fval = _dif_eval(v, thc, jac)
        d = np.sum(jac[1:m + 1] * np.arange(1, m + 1, dtype=np.float64)
                   * np.sin(np.arange(1, m + 1) * thc, dtype=np.float64) * v[1:m + 1])
        thn = thc + fval / d
        if (abs(fval) < tol_bal) or (abs(thc - thn) < np.finfo(np.float64).eps * abs(thc)):
            break
        if fval > 0:
            if (thn < thb) or (thn > tha):
                thn = 0.5 * (thc + tha)
            thb = thc
        else:
            if (thn < thb) or (thn > tha):
                thn = 0.5 * (thc + thb)
            tha = thc
        thc = thn
    mu = np.cos(np.arange(jac.shape[0]) * thc, dtype=np.float64) * jac
    return mu, thc


def cheb_iv(m, mu, xi):
    xi = np.asarray(xi, dtype=np.float64)
    vkm1 = np.zeros(xi.shape[0], dtype=np.float64)
    vk = np.ones(xi.shape[0], dtype=np.float64)
    yi = vk * mu[0]
    for k in range(1, m + 1):
        scal = 1.0 if k == 1 else 2.0
        vkp1 = scal * xi * vk - vkm1
        yi += mu[k] * vkp1
        np.copyto(vkm1, vk)
        np.copyto(vk, vkp1)
    return yi


def interior_params(pol, aa, bb):
    thb = np.arccos(bb)
    tha = np.arccos(aa)
    mu = np.zeros(pol.User_MaximumDegree + 1, dtype=np.float64)
    v = np.zeros(pol.User_MaximumDegree + 1, dtype=np.float64)
    md = pol.User_MinimumDegree
    v[:md] = np.cos(np.arange(md) * thb, dtype=np.float64) - np.cos(np.arange(md) * tha, dtype=np.float64)
    gam = None
    m = md
    for m in range(md, pol.User_MaximumDegree):
        v[m] = np.cos(m * thb, dtype=np.float64) - np.cos(m * tha, dtype=np.float64)
        jac = damping_kernel(m, max_degree=pol.User_MaximumDegree, kind=pol.User_DampingKernel)
        new_mu, new_thc = rootchb(m, v, jac, tha, thb)
        gam = np.cos(new_thc)
        t = cheb_iv(m, new_mu, np.array([gam], dtype=np.float64))
        vals = cheb_iv(m, new_mu, np.array([aa, bb], dtype=np.float64))
        mu = new_mu / t
        if np.all(vals <= t * pol.User_ConvergenceRatio):
            m += 1
            break
    pol.gam = gam
    pol.AdjustedDegree = m - 1
    pol.AdjustedCoefficients = mu
    return pol


def optimize_on_degree(User_MaximumDegree=1000, User_MinimumDegree=10, User_DampingKernel="Jackson",
                       User_ExtremalIntervalDefinition=1e-9, User_WantedInterval=(36.8, 37.2),
                       User_SpectrumBound=(1.0, 759.0), User_ConvergenceRatio=0.1):
    pol = PolParams(User_MaximumDegree=User_MaximumDegree, User_MinimumDegree=User_MinimumDegree,
                    User_DampingKernel=User_DampingKernel,
                    User_ExtremalIntervalDefinition=User_ExtremalIntervalDefinition,
                    User_ConvergenceRatio=User_ConvergenceRatio)
    ext = User_ExtremalIntervalDefinition
    aa = max(User_WantedInterval[0], User_SpectrumBound[0])
    bb = min(User_WantedInterval[1], User_SpectrumBound[1])
    cc = 0.5 * (User_SpectrumBound[1] + User_SpectrumBound[0])
    dd = 0.5 * (User_SpectrumBound[1] - User_SpectrumBound[0])
    aa = max((aa - cc) / dd, -1.0)
    bb = min((bb - cc) / dd, 1.0)
    if aa <= -1.0 + ext:
        aa = -1.0
    if bb >= 1.0 - ext:
        bb = 1.0
    assert not ((aa <= -1.0 + ext) & (bb >= 1.0 - ext)), "ABORTED. Too greedy! Both ends of the wanted interval are close to the extremal. relax either side"
    pol.theta_b = np.arccos(bb, dtype=np.float64)
    pol.theta_a = np.arccos(aa, dtype=np.float64)
    pol.cc = cc
    pol.dd = dd
    pol.AdjustedDampingKernels = damping_kernel(pol.User_MaximumDegree, max_degree=pol.User_MaximumDegree + 1,
                                                kind=pol.User_DampingKernel)
    if (aa <= -1.0 + ext) or (bb >= 1.0 - ext):
        pol = extremal_params(pol, aa, bb)
    else:
        pol = interior_params(pol, aa, bb)
    pol.User_WantedInterval = User_WantedInterval
    pol.User_SpetrumBound = User_SpectrumBound
    return pol


def cheb_av(apply_A, v, pol, return_rho=False):
    """y = sum_k mu_k T_k((A - cc)/dd) v via the three-term recurrence (T2.py:28-92)."""
    mu = pol.AdjustedCoefficients
    dd, cc, m = pol.dd, pol.cc, pol.AdjustedDegree
    jac = pol.AdjustedDampingKernels
    rho = pol.theta_a - pol.theta_b
    vk = v.copy()
    vkm1 = np.zeros_like(v)
    y = v * mu[0]
    for k in range(1, m + 1):
        scal = 1.0 / dd if k == 1 else 2.0 / dd
        w = apply_A(vk)
        vkp1 = scal * (w - cc * vk) - (vkm1 if k > 1 else 0.0)
        y += mu[k] * vkp1
        vkm1 = vk
        vk = vkp1
        rho -= 2 * jac[k] * np.dot(vk, v) * (np.sin(k * pol.theta_b) - np.sin(k * pol.theta_a)) / k
    rho /= np.pi
    if return_rho:
        return y, rho
    return y


def kpm_estimate(apply_A, n, pol, trials, rng):
    """A_Adiag_PolynomialParams_KpmEstimate (T2.py:96-138): int(mean rho * n) + 1."""
    rho = 0.0
    for _ in range(trials):
        v = rng.standard_normal(n)
        v /= np.sqrt(np.sum(v * v))
        _, r = cheb_av(apply_A, v, pol, return_rho=True)
        rho += r
    return int(rho / trials * n) + 1


def optimize_on_memory(apply_A, n, rng, User_MaximumDegree=1000, User_MinimumDegree=10,
                       User_DampingKernel="Jackson", User_ExtremalIntervalDefinition=1e-10,
                       User_WantedInterval=(36.8, 37.2), User_SpectrumBound=(1.0, 759.0),
                       User_DesignatedStart=None, User_WantedNumberEigenvalue=12,
                       User_AffordableMemoryMargin=5, User_NumberKpmTrials=1000,
                       User_ConvergenceRatio=0.1):
    """Bisection on the right end of the wanted interval (T2.py:142-257)."""
    kw = dict(User_MaximumDegree=User_MaximumDegree, User_MinimumDegree=User_MinimumDegree,
              User_DampingKernel=User_DampingKernel,
              User_ExtremalIntervalDefinition=User_ExtremalIntervalDefinition,
              User_SpectrumBound=User_SpectrumBound, User_ConvergenceRatio=User_ConvergenceRatio)
    if User_DesignatedStart is not None:
        pol = optimize_on_degree(User_WantedInterval=(User_WantedInterval[0], User_DesignatedStart), **kw)
        cnt = kpm_estimate(apply_A, n, pol, User_NumberKpmTrials, rng)
        if cnt > User_WantedNumberEigenvalue * 2:
            interval = [User_WantedInterval[0], User_DesignatedStart]
        else:
            interval = [User_WantedInterval[0], User_SpectrumBound[1]]
    else:
        interval = [User_WantedInterval[0], User_SpectrumBound[1]]
    left, right = interval[0], interval[1]
    pol, cnt = None, None
    while right - left >= 1e-8:
        mid = (left + right) / 2
        interval[1] = mid
        pol = optimize_on_degree(User_WantedInterval=interval, **kw)
        cnt = kpm_estimate(apply_A, n, pol, User_NumberKpmTrials, rng)
        if abs(cnt - User_WantedNumberEigenvalue) <= User_AffordableMemoryMargin:
            break
        elif cnt < User_WantedNumberEigenvalue:
            left = mid
        else:
            right = mid
    return pol, cnt, tuple(interval)
