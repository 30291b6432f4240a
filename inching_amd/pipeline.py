"""End-to-end lowest-mode solver: coordinates -> k non-trivial modes at a relative residual.

Protocol (the reference's CTRLM workflow, ``GoogleColab/GoogleColab_Inching_v023_ReleaseOkay.ipynb``
cell 9 and ``Notebook/Application/04_Inching_EigenproblemManyCTRLM.ipynb`` cell 5):
spectrum bound -> Chebyshev filter for the low end [PlusI, a] -> TRLM on p(A) with the
rigid-body block deflated by a Hotelling shift -> Rayleigh quotients with A.

Precision modes:
* ``"f64"``   float64 block values throughout (the plain FP64 GPU solver);
* ``"mixed"`` phase 1 runs the float32 operator (values and Chebyshev vectors float32,
              float64 basis / re-orthogonalisation / Rayleigh-Ritz); phase 2 restarts
              TRLM with a float64 operator (``phase2``: 'mf64' matrix-free or 'f64' stored)
              from a combination of the phase-1 Ritz vectors (an excellent start vector,
              so few float64 restarts are needed); a float32 operator alone stalls at
              ||E|| ~ 6e-8 ||A|| relative residual;
* ``"mf64"``  the matrix-free float64 operator throughout (no stored values).
Acceptance is always the float64 residual ``||A v - lam v|| / (lam_max ||v||)`` computed
with the matrix-free float64 operator, after a float64 Rayleigh-Ritz polish.
"""
import time
import warnings

import numpy as np
import torch

from . import chebyshev as CH
from .operator import Operator
from .polish import apply_fp64, orthonormalise, polish, residual_norms
from .trlm import TRLM


class ModeResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)

    def as_dict(self):
        return {k: v for k, v in self.__dict__.items() if not isinstance(v, torch.Tensor)}


STOP_SAFETY = 0.3   # a cycle ends when the predicted worst relative residual is 0.3 x the target


class NotConvergedWarning(RuntimeWarning):
    """The solver ran out of iterations before every pair met the tolerance; the result holds
    the best pairs found (the reference prints an 'ABORTED. It did not actually converged!'
    line and returns its partial result, ``JacobiDavidsonHotellingDeflation/T1.py:745-757``)."""


def _final(A, Q, lam_max, tol_rel, k, X, state, solver, iters):
    """(theta, X, rel, converged) of the last float64 check of the Ritz rows X [k, n]; when no check ran (the Lanczos
    estimate never came within reach of the target before maxiter), one float64 polish of
    the current Ritz block. Warns when the worst relative residual misses ``tol_rel``."""
    if state["last"] is None:
        theta, Xp, info = polish(A, X[:k], Q=Q, lam_max=lam_max)
        state["last"] = (theta, Xp, info["res"] / lam_max)
    theta, Xp, rel = state["last"]
    converged = bool(np.max(rel) <= tol_rel)
    if not converged:
        warnings.warn("ABORTED. It did not actually converge: %s stopped after %d restarts with a worst relative "
                      "residual of %.3e > %.1e; returning the best pairs found" % (solver, iters, float(np.max(rel)),
                                                                                   tol_rel),
                      NotConvergedWarning, stacklevel=3)
    return theta, Xp, rel, converged


def _check_factory(A, Q, lam_max, tol_rel, k, stall=None):
    """TRLM callback: float64 RR polish of the current Ritz block; stop when all k pairs
    meet ``tol_rel`` or (``stall`` given) when the worst residual improves by less than
    that factor between restarts. Each polish also calibrates the ratio between the worst
    true relative residual and the Lanczos estimate max |beta_k| of the same restart, and
    sets the solver's in-cycle threshold so that a cycle ends once the calibrated estimate
    predicts STOP_SAFETY x tol_rel."""
    state = {"calls": 0, "last": None, "hist": [], "armed": False}

    def check(trlm, S):
        state["calls"] += 1
        # the float64 check costs a k-vector polish; run it once the Lanczos residual
        # estimate is within reach of the target (then every restart), else every 4th
        gate = 1e-5 * np.sqrt(k) * max(1.0, float(np.max(np.abs(S))))
        if not state["armed"] and trlm.last_res > gate and state["calls"] % 4 != 0:
            return False
        state["armed"] = True
        theta, X, info = polish(A, trlm.V[:k], Q=Q, lam_max=lam_max)
        rel = info["res"] / lam_max
        worst = float(np.max(rel))
        state["last"] = (theta, X, rel)
        state["hist"].append(worst)
        if worst <= tol_rel:
            return True
        h = state["hist"]
        if stall is not None and len(h) >= 2 and h[-1] > stall * h[-2]:
            return True
        if trlm.last_est and worst > 0.0:
            trlm.stop_est = STOP_SAFETY * tol_rel * trlm.last_est / worst
        return False

    return check, state


def lowest_modes(A, k, Q=None, precision="mixed", tol_rel=1e-8, filtered=True, bound=None, wanted_right=None,
                 max_degree=5120, min_degree=5, ratio=0.7, hotelling=None, factor=3, maxiter=200, seed=0,
                 phase1_rel=3e-7, v0=None, verbose=False, phase2="mf64", timer=None, kpm_count=None):
    """k lowest non-trivial eigenpairs of A = H + PlusI I (RCM order) on the device.

    ``factor`` sets the Lanczos basis to factor * k vectors (the reference's
    User_WorkspaceSizeFactor); 3 is the default here because with restart cycles that end
    once the residual estimates predict the target, fewer, longer cycles took 6-8% less
    time than factor 2 on 300k-atom globules (64 modes, two seeds).

    Returns ModeResult with eigval (ascending, includes PlusI), eigvec rows [k, n] (float64
    device), rel_res, per-phase seconds and SpMV counts."""
    t_all = time.time()
    phases = {}
    n_spmv = {}
    torch.cuda.synchronize()
    t = time.time()
    if bound is None:
        binfo = {}
        lo, hi = CH.spectrum_bound(A, precision="f32" if precision == "mixed" else precision,
                                   stop_rel=CH.BOUND_STOP_REL, info=binfo)
        bound = (A.plus_i, hi * (1 + 1e-3) + 1e-12)
        n_spmv["bound"] = binfo["products"]
    lam_max = bound[1]
    phases["bound"] = time.time() - t
    pol = None
    if filtered:
        t = time.time()
        if wanted_right is None:
            # KPM bisection for an interval holding about 3 (k + rigid) eigenvalues, as in the
            # reference's CTRLM set-up (WantedNumber = (k - 5) * 3, margin 25, 5 probes)
            pol, cnt, interval, nk = CH.optimize_on_memory(
                A, bound, A.plus_i, wanted_count=int(kpm_count) if kpm_count else max(3 * k, k + 32),
                designated_start=bound[1] / 10.0,
                max_degree=max_degree, min_degree=min_degree, margin=max(5, k // 2), trials=4, ratio=ratio,
                seed=seed, precision="f32" if precision == "mixed" else precision)
            n_spmv["kpm"] = nk
        else:
            pol = CH.design_filter(A.plus_i, wanted_right, bound, max_degree, min_degree, ratio)
        phases["filter"] = time.time() - t
    if hotelling is None:
        hotelling = -40.0 if filtered else 10.0
    check, state = _check_factory(A, Q, lam_max, tol_rel, k)
    start = v0
    total = 0
    applies = 0
    if precision == "mixed":
        t = time.time()
        op32 = Operator(A, "f32", pol=pol, Q=Q, shift=hotelling, timer=timer)
        # phase 1: float32 operator until the float64-verified residual reaches phase1_rel
        # or stops improving (the float32 floor)
        check1, state1 = _check_factory(A, Q, lam_max, phase1_rel, k, stall=0.7)
        r1 = TRLM(op32, k, factor=factor).run(maxiter=maxiter, tol=0.0, v0=start, seed=seed, filtered=filtered,
                                                check=check1, verbose=verbose,
                                                stop_est=STOP_SAFETY * phase1_rel * (1.0 if filtered else lam_max))
        phases["phase1_f32"] = time.time() - t
        n_spmv["phase1_f32"] = op32.n_spmv
        total += op32.n_spmv
        applies += op32.n_apply
        if state1["last"] is None:
            theta1, X1, info1 = polish(A, r1.eigvec, Q=Q, lam_max=lam_max)
            state1["last"] = (theta1, X1, info1["res"] / lam_max)
        theta1, X1, rel1 = state1["last"]
        if float(np.max(rel1)) <= tol_rel:
            torch.cuda.synchronize()
            return ModeResult(eigval=np.asarray(theta1), eigvec=X1, rel_res=np.asarray(rel1), lam_max=lam_max,
                              phases=phases, n_spmv=n_spmv, total_spmv=total, restarts=r1.iters + 1, n_apply=applies,
                              ncv=min(max(factor * k, k + 32), A.shape[0] - 1),
                              degree=int(pol.AdjustedDegree) if pol is not None else 0,
                              seconds=time.time() - t_all, precision=precision, converged=True)
        g = torch.Generator(device=A.device)
        g.manual_seed(seed + 1)
        c = torch.rand(k, generator=g, dtype=torch.float64, device=A.device) + 0.5
        start = (c[:, None] * X1).sum(0)
        del op32, r1
    t = time.time()
    p2 = phase2 if precision == "mixed" else precision
    op64 = Operator(A, p2, pol=pol, Q=Q, shift=hotelling, timer=timer if precision != "mixed" else None)
    # until a polish calibrates it, the in-cycle stop assumes the Lanczos estimate of the
    # filtered operator equals the relative residual (measured: it is about 3x larger)
    r2 = TRLM(op64, k, factor=factor).run(maxiter=maxiter, tol=0.0, v0=start, seed=seed, filtered=filtered,
                                           check=check, verbose=verbose,
                                           stop_est=STOP_SAFETY * tol_rel * (1.0 if filtered else lam_max))
    key = ("phase2_" + p2) if precision == "mixed" else ("solve_" + p2)
    phases[key] = time.time() - t
    n_spmv[key] = op64.n_spmv
    total += op64.n_spmv
    applies += op64.n_apply
    theta, X, rel, converged = _final(A, Q, lam_max, tol_rel, k, r2.eigvec, state, "TRLM", r2.iters + 1)
    torch.cuda.synchronize()
    return ModeResult(eigval=np.asarray(theta), eigvec=X, rel_res=np.asarray(rel), lam_max=lam_max,
                      phases=phases, n_spmv=n_spmv, total_spmv=total, restarts=r2.iters + 1, n_apply=applies,
                      ncv=min(max(factor * k, k + 32), A.shape[0] - 1),
                      degree=int(pol.AdjustedDegree) if pol is not None else 0,
                      seconds=time.time() - t_all, precision=precision, converged=converged)


def lowest_modes_warm(A, k, net, Q=None, precision="mf64", tol_rel=1e-8, **kw):
    """lowest_modes started from the learned warm block: Rayleigh-Ritz of the predicted
    block (float64, true operator), then TRLM from the combination of its k lowest Ritz
    vectors. Falls back to the cold start when the warm solve misses the tolerance, so
    the returned modes never depend on the network."""
    from .warmstart import warm_block
    t0 = time.time()
    lo, hi = CH.spectrum_bound(A, precision="f32" if precision == "mixed" else precision, stop_rel=CH.BOUND_STOP_REL)
    bound = (A.plus_i, hi * (1 + 1e-3) + 1e-12)
    theta, X, info = warm_block(net, A, Q=Q, lam_max=bound[1])
    t_warm = time.time() - t0
    if not _usable_block(theta, info):
        return _cold_fallback(lowest_modes, A, k, Q, precision, tol_rel, bound, t0, kw)
    m = min(k, X.shape[0])
    g = torch.Generator(device=A.device)
    g.manual_seed(int(kw.get("seed", 0)) + 7)
    c = torch.rand(m, generator=g, dtype=torch.float64, device=A.device) + 0.5
    v0 = (c[:, None] * X[:m]).sum(0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", NotConvergedWarning)     # the cold start below decides
        r = lowest_modes(A, k, Q=Q, precision=precision, tol_rel=tol_rel, bound=bound, v0=v0, **kw)
    r.phases["warm_block"] = t_warm
    r.seconds = time.time() - t0
    r.warm_rel_res = info.get("rel_res")
    r.fallback = False
    if not r.converged:
        return _cold_fallback(lowest_modes, A, k, Q, precision, tol_rel, bound, t0, kw)
    return r


def _usable_block(theta, info):
    """A warm block is used only when its Rayleigh-Ritz values and residuals are finite (a
    network with corrupted weights predicts NaN or Inf)."""
    return bool(np.all(np.isfinite(theta)) and np.all(np.isfinite(info["res"])))


def _cold_fallback(solver, A, k, Q, precision, tol_rel, bound, t0, kw):
    cold = solver(A, k, Q=Q, precision=precision, tol_rel=tol_rel, bound=bound, **kw)
    cold.fallback = True
    cold.seconds = time.time() - t0
    return cold


def lowest_modes_jdm(A, k, Q=None, precision="mf64", tol_rel=1e-8, V0=None, shift=40.0, factor=2, maxiter=5000,
                     seed=0, bound=None):
    """k lowest non-trivial eigenpairs by Jacobi-Davidson on A (+ shift Q^T Q), the method of
    ``Burn/JacobiDavidsonHotellingDeflation/T1.py:382-757``, optionally started from the
    search space V0 (rows). Accepted on the float64 residual after a float64 Rayleigh-Ritz
    polish, like ``lowest_modes``."""
    from .jdm import JDM
    if precision not in ("f64", "mf64", "mixed"):
        raise ValueError("precision must be 'f64', 'mf64' or 'mixed'")
    t0 = time.time()
    n_spmv = {}
    mixed = precision == "mixed"
    if bound is None:
        binfo = {}
        lo, hi = CH.spectrum_bound(A, precision="f32" if mixed else precision, stop_rel=CH.BOUND_STOP_REL,
                                   info=binfo)
        bound = (A.plus_i, hi * (1 + 1e-3) + 1e-12)
        n_spmv["bound"] = binfo["products"]
    lam_max = bound[1]
    op = Operator(A, "mf64" if mixed else precision, Q=Q, shift=shift)
    inner = Operator(A, "f32", Q=Q, shift=shift) if mixed else None
    t = time.time()
    r = JDM(op, k, factor=factor, inner=inner).run(tol=0.5 * tol_rel * lam_max, maxiter=maxiter, V0=V0, seed=seed)
    phases = {"solve_jdm_" + precision: time.time() - t}
    n_spmv["jdm"] = op.n_spmv
    if mixed:
        n_spmv["jdm_correction_f32"] = inner.n_spmv
    theta, X, info = polish(A, r.eigvec, Q=Q, lam_max=lam_max)
    state = {"last": (theta, X, info["rel_res"])}
    theta, X, rel, converged = _final(A, Q, lam_max, tol_rel, k, X, state, "JDM", int(r.iters))
    torch.cuda.synchronize()
    return ModeResult(eigval=np.asarray(theta), eigvec=X, rel_res=np.asarray(rel), lam_max=lam_max,
                      phases=phases, n_spmv=n_spmv, total_spmv=r.n_spmv, restarts=int(r.iters),
                      n_apply=op.n_apply + (inner.n_apply if mixed else 0),
                      ncv=factor * k, degree=0, seconds=time.time() - t0, precision=precision,
                      n_converged=int(r.converged), converged=converged and int(r.converged) >= k)


def lowest_modes_jdm_warm(A, k, net, Q=None, precision="mf64", tol_rel=1e-8, **kw):
    """lowest_modes_jdm started from the learned warm block (its float64 Ritz vectors as the
    initial search space). Falls back to the cold start when the warm solve misses the
    tolerance, so the returned modes never depend on the network."""
    from .warmstart import warm_block
    t0 = time.time()
    lo, hi = CH.spectrum_bound(A, precision="f32" if precision == "mixed" else precision, stop_rel=CH.BOUND_STOP_REL)
    bound = (A.plus_i, hi * (1 + 1e-3) + 1e-12)
    theta, X, info = warm_block(net, A, Q=Q, lam_max=bound[1])
    if not _usable_block(theta, info):
        return _cold_fallback(lowest_modes_jdm, A, k, Q, precision, tol_rel, bound, t0, kw)
    factor = int(kw.get("factor", 2))
    V0 = X[:min(X.shape[0], factor * k - 1)].contiguous()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", NotConvergedWarning)     # the cold start below decides
        r = lowest_modes_jdm(A, k, Q=Q, precision=precision, tol_rel=tol_rel, V0=V0, bound=bound, **kw)
    r.phases["warm_block"] = time.time() - t0 - r.seconds
    r.seconds = time.time() - t0
    r.warm_rel_res = info.get("rel_res")
    r.fallback = False
    if not r.converged:
        return _cold_fallback(lowest_modes_jdm, A, k, Q, precision, tol_rel, bound, t0, kw)
    return r
