"""Chebyshev polynomial filters, KPM eigenvalue counts and the spectrum bound.

Filter design mirrors ``Burn/PolynomialFilters/T0.py`` (PolParams :4-20, damping kernels
:25-55, extremal filter :64-169, balancing Newton/bisection :172-226, ChebIv :229-245,
interior filter :248-303, OOC0_OptimizePolynomialParamsOnDegree :313-395); it runs on the
host in float64 and reproduces the reference's coefficient vectors bit for bit (the
arithmetic order is kept; checked against vectors written by the reference's module).

Device parts:

* ``cheb_av``          filtered matvec (``T2.py:28-92``): one fused SpMV + three-term
                       recurrence + accumulate kernel per degree (operator.Operator).
* ``KPMMoments``       the KPM count of ``A_Adiag_PolynomialParams_KpmEstimate``
                       (``T2.py:96-138``) needs the moments mu_k = v^T T_k(A') v of the
                       probe vectors; they do not depend on the filter, so they are
                       computed once (all probes as one block, SpMM per degree) and reused
                       by every bisection step of ``optimize_on_memory`` (``T2.py:142-257``),
                       instead of re-running the recurrence per step.
* ``spectrum_bound``   ``Burn/HermitianLanczos/T2.py:30-101``: 300-step Lanczos without
                       re-orthogonalisation; device vectors, host eigvalsh of T.
"""
import math

import numpy as np
import torch

from . import ops


class PolParams:
    def __init__(self, User_MaximumDegree=10000, User_MinimumDegree=2, User_DampingKernel="Jackson", mu=None, deg=0,
                 User_ExtremalIntervalDefinition=1e-9, User_ConvergenceRatio=0.1):
        self.User_MaximumDegree = User_MaximumDegree
        self.User_MinimumDegree = User_MinimumDegree
        self.User_DampingKernel = User_DampingKernel
        self.User_ExtremalIntervalDefinition = User_ExtremalIntervalDefinition
        self.User_ConvergenceRatio = User_ConvergenceRatio
        self.AdjustedCoefficients = mu
        self.AdjustedDampingKernels = None
        self.AdjustedDegree = deg


def OOC1_DegreeM_DampingKernel(m, User_MaximumDegree=1000, User_DampingKernel="Jackson", dtype_temp=np.float64):
    """Damping factors g_k, k = 0..m (Jackson / Lanczos sigma / none), g_0 = 1/2."""
    g = np.zeros(User_MaximumDegree + 1, dtype=dtype_temp)
    kk = np.arange(1, m + 1, dtype=np.float64)
    if User_DampingKernel == "Jackson":
        t = np.pi / (m + 2)
        inv = 1 / (m + 2)
        g[1:m + 1] = inv * np.sin((kk + 1) * t, dtype=dtype_temp) / np.sin(t, dtype=dtype_temp) + \
            (1 - (kk + 1) * inv) * np.cos(kk * t, dtype=dtype_temp)
    elif User_DampingKernel == "LanczosSigma":
        t = np.pi / (m + 1)
        g[1:m + 1] = np.sin(kk * t, dtype=dtype_temp) / (kk * t)
    else:
        g[1:] = 1.0
    g[0] = 0.5
    return g


def _extremal(pol, a, b):
    """Extremal (one-sided) filter: scaled Chebyshev recurrence, degree where it drops
    below 0.1 at the interval end (and >= minimum degree), coefficients x damping."""
    gam = np.sign(b - (1 - pol.User_ExtremalIntervalDefinition))
    x = a if gam > 0 else b
    a, b = tuple(sorted([x - gam * 0.1 * np.sqrt((1 - gam * x) / 2, dtype=np.float64), -1 * gam]))
    c = (b + a) * 0.5
    e = (b - a) * 0.5
    sig1 = e / (gam - c)
    # 1) degree by the scalar recurrence at x
    s = sig1
    p0, p1 = 1.0, (x - c) * (sig1 / e)
    deg = 1
    for k in range(2, pol.User_MaximumDegree + 1):
        s_new = 1.0 / (2.0 / sig1 - s)
        r = s_new / e
        p0, p1 = np.copy(p1), np.copy(2 * (x - c) * r * p1 - (s * s_new) * p0)
        s = s_new
        deg += 1
        if (p1 < 0.1) & (k >= pol.User_MinimumDegree):
            break
    bar = np.copy(p1)
    # 2) the same recurrence on Chebyshev coefficient vectors
    md = pol.User_MaximumDegree
    s = sig1
    tau = 2 / sig1
    c0 = np.zeros(md + 1, dtype=np.float64)
    c1 = np.zeros(md + 1, dtype=np.float64)
    cn = np.zeros(md + 1, dtype=np.float64)
    c0[0] = 1.0
    c1[0] = -c * (sig1 / e)
    c1[1] = 1 * (sig1 / e)
    for k in range(2, deg + 1):
        s_new = 1.0 / (tau - s)
        r = s_new / e
        cn[:k] = c1[1:k + 1]
        cn[1] += c1[0]
        cn[1:k + 1] += c1[:k]
        cn *= r
        cn[:k + 1] -= (2 * r * c) * c1[:k + 1]
        cn[:k + 1] -= (s * s_new) * c0[:k + 1]
        c0[:k + 1] = np.copy(c1[:k + 1])
        c1[:k + 1] = np.copy(cn[:k + 1])
        s = s_new
    pol.AdjustedCoefficients = c1[:deg + 1] * pol.AdjustedDampingKernels[:deg + 1]
    pol.AdjustedDegree = deg
    pol.bar = bar
    pol.gam = gam
    return pol


def _balance(m, v, g, tha, thb):
    """Angle thc where the damped expansion of the interval indicator balances
    (Newton with bisection safeguard, at most 759 steps)."""
    tol = abs(tha - thb) * 1e-6
    thc = (tha + thb) / 2
    idx_all = np.arange(g.shape[0])
    idx_m = np.arange(1, m + 1)
    idx_mf = np.arange(1, m + 1, dtype=np.float64)
    eps = np.finfo(np.float64).eps
    for _ in range(759):
        f = np.sum(v * np.cos(idx_all * thc, dtype=np.float64) * g)
        df = np.sum(g[1:m + 1] * idx_mf * np.sin(idx_m * thc, dtype=np.float64) * v[1:m + 1])
        th_new = thc + f / df
        if abs(f) < tol or abs(thc - th_new) < eps * abs(thc):
            break
        if f > 0:
            if th_new < thb or th_new > tha:
                th_new = 0.5 * (thc + tha)
            thb = thc
        else:
            if th_new < thb or th_new > tha:
                th_new = 0.5 * (thc + thb)
            tha = thc
        thc = th_new
    return np.cos(idx_all * thc, dtype=np.float64) * g, thc


def ChebIv(m, mu, xi, dtype_temp=np.float64):
    """sum_k mu_k T_k(xi) by the three-term recurrence."""
    xi = np.asarray(xi, dtype=dtype_temp)
    t_prev = np.zeros(xi.shape[0], dtype=dtype_temp)
    t_cur = np.ones(xi.shape[0], dtype=dtype_temp)
    y = t_cur * mu[0]
    for k in range(1, m + 1):
        t_next = (1.0 if k == 1 else 2.0) * xi * t_cur - t_prev
        y += mu[k] * t_next
        np.copyto(t_prev, t_cur)
        np.copyto(t_cur, t_next)
    return y


def _interior(pol, aa, bb):
    thb = np.arccos(bb)
    tha = np.arccos(aa)
    md = pol.User_MaximumDegree
    mind = pol.User_MinimumDegree
    mu = np.zeros(md + 1, dtype=np.float64)
    v = np.zeros(md + 1, dtype=np.float64)
    v[:mind] = np.cos(np.arange(mind) * thb, dtype=np.float64) - np.cos(np.arange(mind) * tha, dtype=np.float64)
    gam = None
    m = mind
    for m in range(mind, md):
        v[m] = np.cos(m * thb, dtype=np.float64) - np.cos(m * tha, dtype=np.float64)
        g = OOC1_DegreeM_DampingKernel(m, User_MaximumDegree=md, User_DampingKernel=pol.User_DampingKernel)
        cand, thc = _balance(m, v, g, tha, thb)
        gam = np.cos(thc)
        t = ChebIv(m, cand, np.array([gam], dtype=np.float64))
        ends = ChebIv(m, cand, np.array([aa, bb], dtype=np.float64))
        mu = cand / t
        if np.all(ends <= t * pol.User_ConvergenceRatio):
            m += 1
            break
    pol.gam = gam
    pol.AdjustedDegree = m - 1
    pol.AdjustedCoefficients = mu
    return pol


def OOC0_OptimizePolynomialParamsOnDegree(User_MaximumDegree=1000, User_MinimumDegree=10, User_DampingKernel="Jackson",
                                          User_ExtremalIntervalDefinition=1e-9, User_WantedInterval=(36.8, 37.2),
                                          User_SpectrumBound=(1.0, 759.0), User_ConvergenceRatio=0.1):
    pol = PolParams(User_MaximumDegree=User_MaximumDegree, User_MinimumDegree=User_MinimumDegree,
                    User_DampingKernel=User_DampingKernel,
                    User_ExtremalIntervalDefinition=User_ExtremalIntervalDefinition,
                    User_ConvergenceRatio=User_ConvergenceRatio)
    ext = User_ExtremalIntervalDefinition
    lo, hi = User_SpectrumBound
    cc = 0.5 * (hi + lo)
    dd = 0.5 * (hi - lo)
    aa = max((max(User_WantedInterval[0], lo) - cc) / dd, -1.0)
    bb = min((min(User_WantedInterval[1], hi) - cc) / dd, 1.0)
    if aa <= -1.0 + ext:
        aa = -1.0
    if bb >= 1.0 - ext:
        bb = 1.0
    assert not ((aa <= -1.0 + ext) & (bb >= 1.0 - ext)), \
        "ABORTED. Too greedy! Both ends of the wanted interval are close to the extremal. relax either side"
    pol.theta_b = np.arccos(bb, dtype=np.float64)
    pol.theta_a = np.arccos(aa, dtype=np.float64)
    pol.cc = cc
    pol.dd = dd
    pol.AdjustedDampingKernels = OOC1_DegreeM_DampingKernel(User_MaximumDegree, User_MaximumDegree=User_MaximumDegree + 1,
                                                            User_DampingKernel=User_DampingKernel)
    if (aa <= -1.0 + ext) or (bb >= 1.0 - ext):
        pol = _extremal(pol, aa, bb)
    else:
        pol = _interior(pol, aa, bb)
    pol.User_WantedInterval = User_WantedInterval
    pol.User_SpetrumBound = User_SpectrumBound
    return pol


def design_filter(a, b, bound, max_degree=5120, min_degree=5, ratio=0.7, ext=1e-10, kernel="Jackson"):
    return OOC0_OptimizePolynomialParamsOnDegree(User_MaximumDegree=max_degree, User_MinimumDegree=min_degree,
                                                 User_DampingKernel=kernel, User_ExtremalIntervalDefinition=ext,
                                                 User_WantedInterval=(a, b), User_SpectrumBound=bound,
                                                 User_ConvergenceRatio=ratio)


# ------------------------------------------------------------------ device ----

def kpm_count_from_moments(pol, moments, n):
    """The reference's KPM count formula (T2.py:66-80, :118-133) evaluated from the probe
    moments mom[k] = mean_probes v^T T_k(A') v (k >= 1), for the filter ``pol``."""
    m = int(pol.AdjustedDegree)
    jac = pol.AdjustedDampingKernels
    k = np.arange(1, m + 1, dtype=np.float64)
    terms = 2 * jac[1:m + 1] * moments[1:m + 1] * (np.sin(k * pol.theta_b) - np.sin(k * pol.theta_a)) / k
    rho = pol.theta_a - pol.theta_b
    for t in terms:          # the reference accumulates term by term
        rho -= t
    rho /= np.pi
    return int(rho * n) + 1


class KPMMoments:
    """Moments v^T T_k((A - cc)/dd) v averaged over ``trials`` unit Gaussian probes, grown on
    demand; A' depends only on the spectrum bound (cc, dd), not on the wanted interval."""

    def __init__(self, A, bound, trials=5, seed=0, precision="f32"):
        if precision != "mf64":
            A.ensure(precision)
        self.A = A
        self.precision = precision
        lo, hi = bound
        self.cc = 0.5 * (hi + lo)
        self.dd = 0.5 * (hi - lo)
        n = A.shape[0]
        self.n = n
        dev = A.device
        vdt = torch.float32 if precision == "f32" else torch.float64
        g = torch.Generator(device=dev)
        g.manual_seed(int(seed))
        P = torch.randn((n, trials), generator=g, dtype=torch.float64, device=dev)
        P /= torch.linalg.norm(P, dim=0, keepdim=True)
        self.P64 = P
        self.tprev = None
        self.tcur = P.to(vdt).contiguous()
        self.mom = [1.0]
        self.n_spmv = 0
        self.trials = int(trials)
        self._w64 = torch.empty(n * trials, dtype=torch.float64, device=dev)
        self._dot = torch.empty(1, dtype=torch.float64, device=dev)
        self._ws = ops.Workspace(1, dev)

    def _step(self):
        k = len(self.mom)
        if self.precision == "mf64":
            W = self.A.matfree(self.tcur)
        else:
            W = self.A.matmat(self.tcur, precision=self.precision)
        self.n_spmv += self.tcur.shape[1]
        scal = (1.0 if k == 1 else 2.0) / self.dd
        # device elementwise recurrence on the (n x trials) block
        ops.axpby(-self.cc, self.tcur.reshape(-1), 1.0, W.reshape(-1))
        ops.axpby(0.0, W.reshape(-1), scal, W.reshape(-1))
        if self.tprev is not None:
            ops.axpby(-1.0, self.tprev.reshape(-1), 1.0, W.reshape(-1))
        self.tprev, self.tcur = self.tcur, W
        w64 = self._w64
        ops.convert(W.reshape(-1), w64)
        ops.multidot(w64.view(1, -1), 1, self.P64.reshape(-1), self._dot, self._ws)
        self.mom.append(self._dot.item() / self.trials)

    def moments(self, m):
        while len(self.mom) <= m:
            self._step()
        return np.asarray(self.mom[:m + 1])


def optimize_on_memory(A, bound, wanted_left, wanted_count, designated_start=None, max_degree=5120, min_degree=5,
                       kernel="Jackson", ext=1e-10, margin=5, trials=5, ratio=0.1, seed=0, precision="f32"):
    """Bisection of the right end of the wanted interval until the KPM count is within
    ``margin`` of ``wanted_count`` (A_Adiag_OptimizePolynomialParamsOnMemory)."""
    kpm = KPMMoments(A, bound, trials=trials, seed=seed, precision=precision)
    n = A.shape[0]

    def count(pol):
        return kpm_count_from_moments(pol, kpm.moments(pol.AdjustedDegree), n)

    def design(right):
        return design_filter(wanted_left, right, bound, max_degree, min_degree, ratio, ext, kernel)

    if designated_start is not None:
        pol = design(designated_start)
        cnt = count(pol)
        interval = [wanted_left, designated_start] if cnt > wanted_count * 2 else [wanted_left, bound[1]]
    else:
        interval = [wanted_left, bound[1]]
    left, right = interval
    pol, cnt = None, None
    while right - left >= 1e-8:
        mid = (left + right) / 2
        interval[1] = mid
        pol = design(mid)
        cnt = count(pol)
        if abs(cnt - wanted_count) <= margin:
            break
        if cnt < wanted_count:
            left = mid
        else:
            right = mid
    return pol, cnt, tuple(interval), kpm.n_spmv


def jackson_kernel(M):
    """Jackson damping factors g_0..g_M of a degree-M Chebyshev expansion (Weisse, Wellein,
    Alvermann and Fehske, Rev. Mod. Phys. 78, 275 (2006), eq. 71): the damped expansion of a
    non-negative function stays non-negative, with a resolution of about pi / M in the angle
    arccos(x)."""
    q = np.pi / (M + 1)
    k = np.arange(M + 1, dtype=np.float64)
    return ((M - k + 1) * np.cos(k * q) + np.sin(k * q) / np.tan(q)) / (M + 1)


class KPMCounter:
    """Number of eigenvalues of A in [bound[0], r] for any r, from the Chebyshev moments
    mu_k = v^T T_k(A') v of ``probes`` random unit vectors (A' = (A - cc) / dd mapping the
    bound onto [-1, 1]), with the Jackson damping of the expansion degree ``degree``:

        count(r) = n (c_0 + sum_k g_k c_k mu_k),  c_0 = (pi - t) / pi,
        c_k = -2 sin(k t) / (k pi),  t = arccos((r - cc) / dd).

    Unlike the reference's estimate (which damps with the filter's kernel, ~1 for low
    degrees, so narrow intervals at the end of a large spectrum give oscillating, even
    negative, counts), this one is monotone in r and non-negative; its resolution near the
    lower end of the spectrum is about dd (pi / degree)^2 / 2 in eigenvalue. The moments
    come from degree / 2 operator launches (fused Chebyshev steps) by the doubling relations
    mu_2k = 2 <T_k v, T_k v> - mu_0 and mu_2k-1 = 2 <T_k v, T_k-1 v> - mu_1 (two dots per
    step; an odd degree takes one more launch). With the matrix-free operator the recurrence
    runs on the cluster-order records (the probe is drawn in the matrix's order, as otherwise,
    and packed) and the dots on their slot-order outputs, whose padding slots are zero."""

    def __init__(self, A, bound, degree=64, probes=1, seed=0, precision="mf64"):
        if precision != "mf64":
            A.ensure(precision)
        lo, hi = bound
        self.cc, self.dd = 0.5 * (hi + lo), 0.5 * (hi - lo)
        self.n = A.shape[0]
        self.degree = int(degree)
        dev = A.device
        half = (self.degree + 1) // 2
        dots = torch.zeros((probes, 2 * half + 1), dtype=torch.float64, device=dev)
        g = torch.Generator(device=dev)
        g.manual_seed(int(seed))
        ws = ops.Workspace(1, dev)
        rec = A.records() if precision == "mf64" and getattr(A, "images", None) is None else None
        if rec is not None:
            self._dots_on_records(A, rec, g, probes, half, dots, ws)
        else:
            self._dots(A, precision, g, probes, half, dots, ws)
        dh = dots.cpu().numpy()
        mom = np.empty((probes, self.degree + 1))
        mom[:, 0] = 1.0
        mom[:, 1] = dh[:, 1]
        for k in range(1, half + 1):
            if 2 * k <= self.degree:
                mom[:, 2 * k] = 2.0 * dh[:, 2 * k] - 1.0
            if 2 <= 2 * k - 1 <= self.degree:
                mom[:, 2 * k - 1] = 2.0 * dh[:, 2 * k - 1] - mom[:, 1]
        self.moments = mom.mean(0)
        self.jackson = jackson_kernel(self.degree)
        self.n_spmv = probes * half

    def _dots_on_records(self, A, rec, g, probes, half, dots, ws):
        """dots[p, 2k] = <T_k v, T_k v>, dots[p, 2k - 1] = <T_k v, T_k-1 v> for probe p, the
        recurrence on three record buffers in turn, T_k v in slot order from each step."""
        recs, rent = rec
        atoms, ptr = A.clusters()[:2]
        ncl = ptr.numel() - 1
        ys = [torch.empty(12 * ncl, dtype=torch.float64, device=A.device) for _ in range(2)]
        for p in range(probes):
            v = torch.randn(self.n, generator=g, dtype=torch.float64, device=A.device)
            v /= torch.linalg.norm(v)
            ops.rec_pack(atoms, v, recs[0], acc=ys[0], mu0=1.0)          # T_0 v: records and slot order
            prev, cur, free = None, recs[0], [recs[1], recs[2]]
            for k in range(1, half + 1):
                nxt = free.pop(0)
                y, y_prev = ys[k % 2], ys[(k - 1) % 2]
                if k == 1:
                    ops.matfree_rec(cur, None, nxt, ptr, rent, A.plus_i, s1=1.0 / self.dd, c0=self.cc, y=y)
                else:
                    ops.matfree_rec(cur, prev, nxt, ptr, rent, A.plus_i, s1=2.0 / self.dd, c0=self.cc, c2=1.0, y=y)
                ops.multidot(y.view(1, -1), 1, y, dots[p, 2 * k:2 * k + 1], ws)
                ops.multidot(y_prev.view(1, -1), 1, y, dots[p, 2 * k - 1:2 * k], ws)
                if prev is not None:
                    free.append(prev)
                prev, cur = cur, nxt

    def _dots(self, A, precision, g, probes, half, dots, ws):
        """The same dots with the recurrence in the matrix's order (any operator precision)."""
        from .autotune import best_config, bsr_launch
        dev = A.device
        vdt = torch.float32 if precision == "f32" else torch.float64
        cfg = best_config(A, precision)
        blanes, c16 = bsr_launch(A, cfg) if precision != "mf64" else (0, None)
        vals = A.vals(precision) if precision != "mf64" else None

        def step(x, y, **epi):
            if vals is None:
                A.matfree_spmv(x, y, config=cfg, **epi)
            else:
                ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, vals, x, y, lanes=blanes, col16=c16, **epi)

        t = [torch.empty(self.n, dtype=vdt, device=dev) for _ in range(3)]
        d64 = torch.empty(self.n, dtype=torch.float64, device=dev) if vdt != torch.float64 else None
        d64p = torch.empty(self.n, dtype=torch.float64, device=dev) if vdt != torch.float64 else None
        for p in range(probes):
            v = torch.randn(self.n, generator=g, dtype=torch.float64, device=dev)
            v /= torch.linalg.norm(v)
            t[0].copy_(v)
            prev, cur, nxt = None, t[0], t[1]
            prev64 = v                          # T_{k-1} v in float64
            for k in range(1, half + 1):
                if k == 1:
                    step(cur, nxt, s1=1.0 / self.dd, c0=self.cc)
                else:
                    step(cur, nxt, s1=2.0 / self.dd, c0=self.cc, z=prev, c2=1.0)
                src = nxt
                if d64 is not None:
                    ops.convert(nxt, d64)
                    src = d64
                # dots[2k] = <T_k, T_k>, dots[2k - 1] = <T_k, T_{k-1}>
                ops.multidot(src.view(1, -1), 1, src, dots[p, 2 * k:2 * k + 1], ws)
                ops.multidot(prev64.view(1, -1), 1, src, dots[p, 2 * k - 1:2 * k], ws)
                if d64 is not None:
                    d64, d64p = d64p, d64
                    prev64 = d64p
                else:
                    prev64 = src
                free = prev if prev is not None else t[2]
                prev, cur, nxt = cur, nxt, free

    def count(self, r):
        x = min(1.0, max(-1.0, (float(r) - self.cc) / self.dd))
        t = np.arccos(x)
        k = np.arange(1, self.degree + 1, dtype=np.float64)
        ck = -2.0 * np.sin(k * t) / (k * np.pi)
        return self.n * ((np.pi - t) / np.pi + float(np.sum(self.jackson[1:] * ck * self.moments[1:])))

    def right_for(self, count, lo, hi, tol=1e-6):
        """The r in [lo, hi] whose estimated count is ``count`` (bisection; the count is
        monotone in r)."""
        a, b = float(lo), float(hi)
        while b - a > tol * max(1.0, abs(b)):
            m = 0.5 * (a + b)
            if self.count(m) < count:
                a = m
            else:
                b = m
        return 0.5 * (a + b)


# early end of the bound run for callers that use only its maximum (which they widen by
# 1e-3): on the 2.4M-atom assembly 1e-5 stops after 71 products instead of 101 (1e-10), with
# the maximum 2.2e-8 relative below that of 400 steps. A looser 1e-4 is not safe: on a
# 1,200-atom globule it stopped 2.3e-3 below the largest eigenvalue, outside the widening
BOUND_STOP_REL = 1e-5
BOUND_CHECK_EVERY = 10


def spectrum_bound(A, n_tridiag=300, seed=0, precision="f64", stop_rel=None, info=None):
    """(min, max) Ritz values of a 300-step Lanczos run without re-orthogonalisation
    (``HermitianLanczos/T2.py:30-100``).

    With ``stop_rel`` the run ends early, at a multiple of BOUND_CHECK_EVERY steps, once the
    largest Ritz value moved by at most ``stop_rel`` (relative) over the last
    BOUND_CHECK_EVERY steps. Only the maximum is meant to be used then: the largest Ritz
    value converges first and geometrically (its change over 10 steps bounds its error once
    it is that small), while the minimum may still move. Callers widen the maximum by 1e-3
    relative before designing filters on it. ``info["products"]`` receives the number of
    operator products."""
    if precision != "mf64":
        A.ensure(precision)
    n = A.shape[0]
    dev = A.device
    # with the matrix-free operator the run can stay in the record layout's slot order (the
    # Ritz values do not depend on the numbering): vectors of 12 ncl entries, zero on
    # padding slots (whose rows are zero), one pack into a record per product
    rec = A.records() if precision == "mf64" and getattr(A, "images", None) is None else None
    if rec is not None:
        atoms, ptr = A.clusters()[:2]
        n = 3 * atoms.numel()
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    v = torch.rand(n, generator=g, dtype=torch.float64, device=dev) - 0.5
    if rec is not None:
        v.view(-1, 3)[atoms < 0] = 0.0
    v /= torch.linalg.norm(v)
    f = torch.empty_like(v)
    v0 = torch.empty_like(v)
    alpha = torch.zeros(n_tridiag + 1, dtype=torch.float64, device=dev)
    beta = torch.zeros(n_tridiag, dtype=torch.float64, device=dev)
    ws = ops.Workspace(1, dev)
    vals = A.vals(precision) if precision != "mf64" else None
    row = v.view(1, -1)

    from .autotune import best_config, bsr_launch
    cfg = best_config(A, precision) if rec is None else 0
    blanes, c16 = bsr_launch(A, cfg) if precision != "mf64" else (0, None)

    def av(x, y):
        if rec is not None:
            ops.rec_pack(None, x, rec[0][0], ncl=ptr.numel() - 1)
            ops.matfree_rec(rec[0][0], None, None, ptr, rec[1], A.plus_i, y=y)
        elif vals is None:
            A.matfree_spmv(x, y, config=cfg)
        else:
            ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, vals, x, y, lanes=blanes, col16=c16)

    av(v, f)
    ops.multidot(row, 1, f, alpha[0:1], ws)
    ops.axpy_dev(alpha[0:1], -1.0, v, f)
    m = n_tridiag + 1            # size of the tridiagonal matrix built
    prev = None
    for i in range(n_tridiag):
        ops.gemv_update(row, 0, None, f, 0.0, nrm2_out=beta[i:i + 1], ws=ws)   # beta_i = ||f||
        v0, v = v, v0            # the previous vector is kept by swapping buffers
        ops.scale_div(f, beta[i:i + 1], v)
        row = v.view(1, -1)
        av(v, f)
        ops.axpy_dev(beta[i:i + 1], -1.0, v0, f)
        ops.multidot(row, 1, f, alpha[i + 1:i + 2], ws)
        ops.axpy_dev(alpha[i + 1:i + 2], -1.0, v, f)
        if stop_rel is not None and (i + 1) % BOUND_CHECK_EVERY == 0 and i + 1 < n_tridiag:
            mm = i + 2
            top = float(np.linalg.eigvalsh(_tridiag(alpha[:mm].cpu().numpy(), beta[:mm - 1].cpu().numpy()))[-1])
            if prev is not None and abs(top - prev) <= stop_rel * abs(top):
                m = mm
                break
            prev = top
    if info is not None:
        info["products"] = m
    T = _tridiag(alpha[:m].cpu().numpy(), beta[:m - 1].cpu().numpy())
    w = np.linalg.eigvalsh(T)
    return float(np.min(w)), float(np.max(w))


def _tridiag(a, b):
    return np.diag(a) + np.diag(b, -1) + np.diag(b, 1)
