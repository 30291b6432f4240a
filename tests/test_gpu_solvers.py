"""Device eigensolvers against FP64 dense eigh (the oracle for small cases).

Bars: eigenvalue error <= 1e-9 absolute for float64 operators; for the mixed-precision
(float32 SpMV) operator the FP64 polish brings relative residuals to <= 1e-8 and the
eigenvalue error to <= 1e-7 relative (eigenvalues of A = H + I, the solver output
convention of the reference).
"""
import numpy as np
import pytest
import torch

from oracle import hessian as OH
from tests.helpers import small_globule, subspace_sines

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem(cuda):
    from inching_amd.hessian import build_hessian, rigid_body_basis
    X = small_globule(1200, seed=31)
    A = build_hessian(X, 8.0, plus_i=1.0, reorder=True, precision=("f32", "f64"))
    Xp = X[A.order]
    Ad = OH.hessian_dense(Xp, 8.0, 1.0)
    w, U = np.linalg.eigh(Ad)
    Q = rigid_body_basis(A.xyz)
    return A, Ad, w, U, Q


def _check(Ad, w, U, lam, V, k, tol_val, tol_res):
    V = V / np.linalg.norm(V, axis=0)
    assert np.allclose(lam, w[6:6 + k], rtol=0, atol=tol_val), (lam - w[6:6 + k])
    R = Ad @ V - V * lam[None, :]
    rel = np.linalg.norm(R, axis=0) / w[-1]
    assert rel.max() < tol_res, rel.max()
    assert subspace_sines(U[:, 6:6 + k], V).max() < 1e-3


def test_trlm_f64_hotelling(problem):
    from inching_amd.operator import Operator
    from inching_amd.trlm import TRLM
    A, Ad, w, U, Q = problem
    print("case: trlm f64 hotelling k=16", flush=True)
    k = 16
    op = Operator(A, "f64", Q=Q, shift=10.0)
    r = TRLM(op, k).run(maxiter=500, tol=1e-10)
    lam = r.eigval
    V = r.eigvec.T.cpu().numpy()
    idx = np.argsort(lam)
    _check(Ad, w, U, lam[idx], V[:, idx], k, 1e-9, 1e-10)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("mf64", 1e-10), ("f32", 2e-5)])
def test_operator_filter_equals_spectral_filter(problem, precision, tol):
    """Operator.filtered (the Chebyshev recurrence as fused operator launches) equals
    U p(w) U^T v from dense eigh, p the designed filter polynomial; the in-place call
    (output = input) gives bitwise the out-of-place result."""
    import torch
    from inching_amd import chebyshev as CH
    from inching_amd.operator import Operator
    A, Ad, w, U, Q = problem
    print("case: operator filter %s" % precision, flush=True)
    pol = CH.design_filter(1.0, w[40], (1.0, w[-1] * (1 + 1e-3)), 5120, 5, 0.7)
    m = int(pol.AdjustedDegree)
    p = np.polynomial.chebyshev.chebval((w - pol.cc) / pol.dd, np.asarray(pol.AdjustedCoefficients)[:m + 1])
    v = np.random.default_rng(3).standard_normal(Ad.shape[0])
    ref = U @ (p * (U.T @ v))
    op = Operator(A, precision, pol=pol)
    vd = torch.as_tensor(v, device="cuda")
    u = torch.empty_like(vd)
    op.apply(vd, u)
    got = u.cpu().numpy()
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref), np.linalg.norm(got - ref) / np.linalg.norm(ref)
    x = vd.clone()
    op.apply(x, x)
    assert torch.equal(x, u)


def test_spectrum_bound_early_stop(problem):
    """The bound run with an early stop ends before 300 steps: with stop_rel 1e-10 its maximum
    is within 1e-10 of the 300-step value and of dense eigh's largest eigenvalue; with the
    solvers' BOUND_STOP_REL it is within 1e-4 (well inside the 1e-3 widening the callers
    apply) after no more products; never above the largest eigenvalue by more than
    rounding; the products are counted."""
    from inching_amd import chebyshev as CH
    A, Ad, w, U, Q = problem
    print("case: spectrum bound early stop", flush=True)
    lo_f, hi_f = CH.spectrum_bound(A, precision="f64")
    info = {}
    lo_e, hi_e = CH.spectrum_bound(A, precision="f64", stop_rel=1e-10, info=info)
    assert info["products"] < 301 and info["products"] % CH.BOUND_CHECK_EVERY == 1, info
    assert abs(hi_e - hi_f) <= 1e-10 * hi_f
    assert abs(hi_e - w[-1]) <= 1e-10 * w[-1] and hi_e <= w[-1] * (1 + 1e-14)
    loose = {}
    _, hi_l = CH.spectrum_bound(A, precision="f64", stop_rel=CH.BOUND_STOP_REL, info=loose)
    assert loose["products"] <= info["products"] and loose["products"] % CH.BOUND_CHECK_EVERY == 1, loose
    assert abs(hi_l - w[-1]) <= 1e-4 * w[-1] and hi_l <= w[-1] * (1 + 1e-14), (hi_l, w[-1])
    mf = {}
    _, hi_m = CH.spectrum_bound(A, precision="mf64", stop_rel=1e-10, info=mf)
    assert abs(hi_m - w[-1]) <= 1e-10 * w[-1], (hi_m, w[-1], mf)


def test_trlm_basis_over_256_vectors(problem):
    """k = 90 with ncv = 270: the delayed-CGS2 step past 256 basis vectors (coefficients read
    from global memory) against dense eigh."""
    from inching_amd.operator import Operator
    from inching_amd.trlm import TRLM
    A, Ad, w, U, Q = problem
    print("case: trlm f64 k=90 ncv=270", flush=True)
    k = 90
    op = Operator(A, "f64", Q=Q, shift=40.0)        # the 96th eigenvalue is 12.6: park the rigid modes at 41
    t = TRLM(op, k, factor=3)
    assert t.ncv == 270 and t.dcgs2
    r = t.run(maxiter=500, tol=1e-10)
    lam = r.eigval
    V = r.eigvec.T.cpu().numpy()
    idx = np.argsort(lam)
    _check(Ad, w, U, lam[idx], V[:, idx], k, 1e-9, 1e-10)


def test_mixed_precision_two_phase(problem):
    from inching_amd.pipeline import lowest_modes
    A, Ad, w, U, Q = problem
    print("case: lowest_modes mixed unfiltered k=16", flush=True)
    k = 16
    r = lowest_modes(A, k, Q=Q, precision="mixed", filtered=False, tol_rel=1e-9, bound=(1.0, w[-1] * 1.001))
    V = r.eigvec.T.cpu().numpy()
    _check(Ad, w, U, r.eigval, V, k, 1e-9, 1e-9)
    assert r.n_spmv["phase1_f32"] > 0


def test_mixed_precision_filtered_kpm(problem):
    from inching_amd.pipeline import lowest_modes
    A, Ad, w, U, Q = problem
    print("case: lowest_modes mixed filtered+KPM k=12", flush=True)
    k = 12
    r = lowest_modes(A, k, Q=Q, precision="mixed", filtered=True, tol_rel=1e-8)
    V = r.eigvec.T.cpu().numpy()
    _check(Ad, w, U, r.eigval, V, k, 1e-8, 1e-8)


def test_jdm_mixed_precision(problem):
    """Jacobi-Davidson with the correction solves on float32 stored values (float64 vectors)
    and the outer iteration, residuals and acceptance on the float64 matrix-free operator:
    the same modes as dense eigh to the float64 bars, the float32 operator used, and the
    eigenvalues equal to the all-float64 JDM's."""
    from inching_amd.pipeline import lowest_modes_jdm
    A, Ad, w, U, Q = problem
    print("case: lowest_modes_jdm mixed k=12", flush=True)
    k = 12
    r = lowest_modes_jdm(A, k, Q=Q, precision="mixed", tol_rel=1e-8)
    assert r.converged and np.max(r.rel_res) <= 1e-8
    V = r.eigvec.T.cpu().numpy()
    _check(Ad, w, U, r.eigval, V, k, 1e-8, 1e-8)
    assert r.n_spmv["jdm_correction_f32"] > 0 and r.total_spmv == r.n_spmv["jdm"] + r.n_spmv["jdm_correction_f32"]
    ref = lowest_modes_jdm(A, k, Q=Q, precision="mf64", tol_rel=1e-8)
    assert np.max(np.abs(r.eigval - ref.eigval)) <= 1e-9


def test_trlm_filtered(problem):
    from inching_amd.chebyshev import design_filter
    from inching_amd.operator import Operator
    from inching_amd.trlm import TRLM, rayleigh_quotients
    A, Ad, w, U, Q = problem
    print("case: trlm filtered k=12", flush=True)
    k = 12
    pol = design_filter(1.0, w[6 + k + 8], (1.0, w[-1] + 1e-12), max_degree=2000, min_degree=5, ratio=0.7)
    op = Operator(A, "f64", pol=pol, Q=Q, shift=-40.0)
    r = TRLM(op, k).run(maxiter=500, tol=1e-9, filtered=True)
    rq = rayleigh_quotients(A, r.eigvec).cpu().numpy()
    idx = np.argsort(rq)
    _check(Ad, w, U, rq[idx], r.eigvec.T.cpu().numpy()[:, idx], k, 1e-9, 1e-9)


@pytest.mark.parametrize("designated", [None, "tenth"])
def test_device_kpm_bisection_interval_holds_wanted_count(problem, designated):
    """The device bisection (PolynomialFilters/T2.py:142-257 with KPM moments from the device
    operator) settles on an interval whose exact eigenvalue count (dense eigh) is the wanted
    count within the margin plus the spread of 64 KPM probes."""
    from inching_amd import chebyshev as CH
    A, Ad, w, U, Q = problem
    print("case: device KPM bisection %s" % designated, flush=True)
    bound = (1.0, w[-1] * (1 + 1e-3))
    wanted, margin = 80, 8
    start = bound[1] / 10.0 if designated == "tenth" else None
    pol, cnt, interval, nk = CH.optimize_on_memory(A, bound, 1.0, wanted_count=wanted, designated_start=start,
                                                   max_degree=5120, min_degree=5, margin=margin, trials=64,
                                                   ratio=0.7, seed=3, precision="f64")
    assert abs(cnt - wanted) <= margin and nk > 0
    exact = int(np.sum((w >= 1.0 - 1e-9) & (w <= interval[1])))
    assert abs(exact - wanted) <= margin + 8, (exact, interval, cnt)


def test_kpm_moments_by_doubling_match_the_recurrence(problem):
    """The counter's moments (degree / 2 launches and the doubling relations) equal v^T T_k(A') v
    from the plain three-term recurrence on the dense matrix for the same probe (1e-10), for
    an even and an odd degree."""
    import torch
    from inching_amd import chebyshev as CH
    A, Ad, w, U, Q = problem
    print("case: KPM moments by doubling", flush=True)
    bound = (1.0, w[-1] * (1 + 1e-3))
    for degree in (64, 33):
        c = CH.KPMCounter(A, bound, degree=degree, probes=1, seed=11, precision="mf64")
        g = torch.Generator(device=A.device)
        g.manual_seed(11)
        v = torch.randn(A.shape[0], generator=g, dtype=torch.float64, device=A.device)
        v = (v / torch.linalg.norm(v)).cpu().numpy()
        cc, dd = 0.5 * (bound[1] + bound[0]), 0.5 * (bound[1] - bound[0])
        Ap = (Ad - cc * np.eye(Ad.shape[0])) / dd
        t0, t1 = v, Ap @ v
        ref = [v @ t0, v @ t1]
        for _ in range(2, degree + 1):
            t0, t1 = t1, 2.0 * (Ap @ t1) - t0
            ref.append(v @ t1)
        assert c.n_spmv == (degree + 1) // 2
        assert np.max(np.abs(np.asarray(ref) - c.moments)) <= 1e-10, degree


@pytest.mark.parametrize("precision", ["mf64", "f32"])
def test_kpm_counter_matches_exact_counts(problem, precision):
    """The Jackson-damped KPM count of eigenvalues in [PlusI, r] against the exact count from
    dense eigh (1,200 atoms, 3,600 dof; 16 probes, 64 moments), the bisection for r, and
    monotonicity. Tolerance: 8 + 10 % of the exact count (smoothing of about 0.04 in eigenvalue
    near the lower end plus the probes' statistical error)."""
    from inching_amd import chebyshev as CH
    A, Ad, w, U, Q = problem
    print("case: KPM counter %s" % precision, flush=True)
    bound = (1.0, w[-1] * (1 + 1e-3))
    c = CH.KPMCounter(A, bound, degree=64, probes=16, seed=3, precision=precision)
    prev = -1.0
    for r in (2.0, 4.0, 8.0, 15.0, 30.0):
        exact = int(np.sum(w <= r))
        est = c.count(r)
        assert abs(est - exact) <= 8 + 0.1 * exact, (r, est, exact)
        assert est >= prev
        prev = est
    r = c.right_for(200.0, 1.0, bound[1])
    assert abs(c.count(r) - 200.0) < 1.0
    assert abs(np.sum(w <= r) - 200) <= 8 + 20
