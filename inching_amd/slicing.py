"""Spectrum slicing: many lowest modes in windows of k with constant solver memory.

The reference's CTRLM driver (``Notebook/Application/04_Inching_EigenproblemManyCTRLM.ipynb``
cell 5) walks up the spectrum:

* window 0: extremal Chebyshev filter whose right end is found by the KPM bisection
  (``A_Adiag_OptimizePolynomialParamsOnMemory`` with WantedNumber (k-5)*5, margin 25,
  designated start lambda_max/10), filtered TRLM for k pairs with the six rigid-body modes
  deflated (Hotelling shift -40);
* window w > 0: interior filter on (left, left + 3*guess) re-fitted by the same bisection
  (WantedNumber k-5, margin 5), no deflation; the pairs with eigenvalue > left are kept;
* left <- max(kept) - 1e-10 and guess <- |eig[-1] - eig[0]| of the window.

The reference keeps every pair above ``left`` and so can repeat the top mode of a window
or skip modes when a window's interval holds more than k eigenvalues. Here:

* duplicates: a window's pairs are projected off the previous window's accepted vectors
  (two passes) and the remaining block is re-orthonormalised rank-revealingly (directions
  whose norm collapses were already found) before a float64 Rayleigh-Ritz; degenerate
  partners of a boundary mode survive, repeats do not;
* misses: the filtered TRLM returns the k pairs with the largest filter values; if all k
  lie inside the window's interval the interval may hold more than k eigenvalues, so the
  window is redone on an interval of half the width (``saturation`` retries);
* an interval none of whose eigenvalues appears once the k dominant pairs of the filtered
  operator have settled (rel_res <= empty_tol) is empty: the next window starts past it;
* every accepted pair passes the float64 residual test rel_res <= tol_rel.
"""
import time

import numpy as np
import torch

from . import chebyshev as CH
from . import ops
from .operator import Operator
from .polish import _combine, _gram, apply_fp64, residual_norms
from .trlm import TRLM


class SliceResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def _project_off(X, P):
    """X <- X - (X P^T) P twice (rows)."""
    if P is None or P.shape[0] == 0:
        return X
    for _ in range(2):
        C = ops.gram(X, P)                                   # [k, p]
        Y = torch.empty_like(X)
        ops.rr_combine(P, P.shape[0], C.T.contiguous(), Y)
        ops.axpby(-1.0, Y.view(-1), 1.0, X.view(-1))
    return X


def _rank_revealing_rr(A, X, P, lam_max, keep=0.25):
    """Project X off P, keep the directions with Gram eigenvalue > keep (relative to a
    unit-norm row), float64 Rayleigh-Ritz. Returns (theta, rows, rel_res)."""
    X = _project_off(X.clone(), P)
    G = _gram(X, X)
    G = 0.5 * (G + G.T)
    g, W = np.linalg.eigh(G)
    sel = g > keep
    if not np.any(sel):
        return np.zeros(0), X[:0], np.zeros(0)
    X = _combine(X, W[:, sel] / np.sqrt(g[sel]))
    X = _project_off(X, P)
    G = _gram(X, X)                                          # second pass: exact orthonormality
    R = np.linalg.cholesky(0.5 * (G + G.T)).T
    X = _combine(X, np.linalg.inv(R))
    Y = apply_fp64(A, X)
    Hm = _gram(X, Y)
    theta, Cm = np.linalg.eigh(0.5 * (Hm + Hm.T))
    X = _combine(X, Cm)
    Y = _combine(Y, Cm)
    return theta, X, residual_norms(X, Y, theta) / lam_max


def slice_spectrum(A, total, k=28, Q=None, precision="mf64", tol_rel=1e-8, max_degree=5120, maxiter=300,
                   seed=0, interval_guess=0.25, saturation_retries=3, max_windows=None, empty_tol=1e-6,
                   verbose=False):
    """Lowest ``total`` non-trivial eigenpairs of A = H + PlusI I by windows of k.

    Q: rigid-body rows [6, n] (deflated in window 0; also projected off every window's
    block). Returns SliceResult(eigval [total] ascending, eigvec rows [total, n] float64,
    rel_res, windows (per-window dicts), n_spmv, seconds)."""
    ops.require_gpu()
    if k < 8:
        raise ValueError("k must be >= 8 (windows aim at k - 5 eigenvalues)")
    torch.cuda.synchronize()
    t0 = time.time()
    lo, hi = CH.spectrum_bound(A, precision=precision)
    bound = (A.plus_i, hi * (1 + 1e-3) + 1e-12)
    lam_max = bound[1]
    delta = 1e-10 * lam_max
    vals, vecs, rels, windows = [], [], [], []
    prev = Q.contiguous() if Q is not None else None
    left = A.plus_i
    guess = float(interval_guess)
    n_spmv = 301
    w = 0
    empty = 0
    while sum(v.size for v in vals) < total:
        if max_windows is not None and w >= max_windows:
            break
        width = None
        for attempt in range(saturation_retries + 1):
            if w == 0:
                # extremal window: the filter is monotone, so the k largest filter values are
                # the k lowest eigenvalues whatever the interval holds (no saturation test)
                pol, cnt, interval, nk = CH.optimize_on_memory(
                    A, bound, A.plus_i, wanted_count=(k - 5) * 5, designated_start=bound[1] / 10.0,
                    max_degree=max_degree, min_degree=5, margin=25, trials=5, ratio=0.7, seed=seed + w,
                    precision=precision)
                op = Operator(A, precision, pol=pol, Q=Q, shift=-40.0)
            else:
                start = left + 3.0 * guess if width is None else left + width
                if width is None:
                    pol, cnt, interval, nk = CH.optimize_on_memory(
                        A, bound, left, wanted_count=k - 5, designated_start=min(start, bound[1]),
                        max_degree=max_degree, min_degree=5, margin=max(1, min(5, (k - 5) // 3)), trials=5,
                        ratio=0.7, seed=seed + w,
                        precision=precision)
                else:
                    interval = (left, min(start, bound[1]))
                    pol = CH.design_filter(interval[0], interval[1], bound, max_degree, 5, 0.7)
                    nk = 0
                op = Operator(A, precision, pol=pol, shift=0.0)
            n_spmv += nk
            lo_w, hi_w = float(interval[0]), float(interval[1])
            state = {}

            def check(trlm, S, lo_w=lo_w, hi_w=hi_w, state=state):
                theta, X, rel = _rank_revealing_rr(A, trlm.V[:k], prev, lam_max)
                inside = (theta > lo_w - delta) & (theta <= hi_w)
                state["last"] = (theta, X, rel, inside)
                if np.any(inside):
                    return bool(np.all(rel[inside] <= tol_rel))
                # nothing inside yet: the interval is empty once the k dominant pairs of the
                # filtered operator have settled without one of them entering it
                return bool(rel.size == k and np.all(rel <= empty_tol))

            r = TRLM(op, k).run(maxiter=maxiter, tol=0.0, seed=seed + 17 * w + attempt, filtered=True, check=check)
            n_spmv += op.n_spmv
            theta, X, rel, inside = state["last"]
            saturated = w > 0 and int(np.count_nonzero(inside)) >= k
            if verbose:
                print("window %d attempt %d: interval (%.6f, %.6f] degree %d restarts %d, %d inside, saturated %s; "
                      "theta %s rel %s" % (w, attempt, lo_w, hi_w, pol.AdjustedDegree, r.iters + 1,
                                           int(np.count_nonzero(inside)), saturated, np.round(theta, 5),
                                           np.array2string(rel, precision=1)), flush=True)
            if not saturated:
                break
            width = (hi_w - lo_w) / 2.0
        else:
            raise RuntimeError("window %d stays saturated after %d narrowings" % (w, saturation_retries))
        if not np.any(inside):
            # an empty interval: move past it and widen the next guess
            windows.append({"interval": (lo_w, hi_w), "degree": int(pol.AdjustedDegree), "kept": 0,
                            "restarts": int(r.iters) + 1, "spmv": int(op.n_spmv)})
            empty += 1
            if empty > 20 or hi_w >= bound[1]:
                raise RuntimeError("no eigenvalue found above %.6g (spectrum exhausted?)" % left)
            left = hi_w - delta
            guess = 2.0 * max(guess, hi_w - lo_w)
            w += 1
            continue
        empty = 0
        acc = inside & (rel <= tol_rel)
        if not np.any(acc):
            raise RuntimeError("window %d: no pair in (%.6g, %.6g] met the tolerance after %d restarts"
                               % (w, lo_w, hi_w, int(r.iters) + 1))
        idx = np.nonzero(acc)[0]
        # contiguity: accept from the bottom of the window up to the first unconverged pair
        # inside the interval, so the kept set has no hole
        bad_inside = np.nonzero(inside & (rel > tol_rel))[0]
        if bad_inside.size:
            idx = idx[theta[idx] < theta[bad_inside].min()]
        if idx.size == 0:
            raise RuntimeError("window %d: lowest pair in the interval did not converge" % w)
        Xa = X[torch.as_tensor(idx, device=X.device)].contiguous()
        vals.append(theta[idx])
        vecs.append(Xa)
        rels.append(rel[idx])
        windows.append({"interval": (lo_w, hi_w), "degree": int(pol.AdjustedDegree), "kept": int(idx.size),
                        "restarts": int(r.iters) + 1, "spmv": int(op.n_spmv)})
        prev = Xa if Q is None else torch.cat([Q, Xa], 0).contiguous()
        left = float(theta[idx].max()) - delta
        guess = max(float(theta[idx].max() - theta[idx].min()), 1e-6 * lam_max)
        w += 1
    lam = np.concatenate(vals)
    order = np.argsort(lam, kind="stable")[:total]
    V = torch.cat(vecs, 0)[torch.as_tensor(order, device=A.device)]
    torch.cuda.synchronize()
    return SliceResult(eigval=lam[order], eigvec=V, rel_res=np.concatenate(rels)[order], windows=windows,
                       n_spmv=n_spmv, lam_max=lam_max, seconds=time.time() - t0)
