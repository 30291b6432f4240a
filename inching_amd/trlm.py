"""Thick-restart Lanczos with full (delayed) CGS2 re-orthogonalisation on the device.

Follows ``Burn/ThickRestartLanczosHotellingDeflation/T1.py``:
``S_HeigvalTRLMHD_HeigvecTRLMHD`` :60-310 (driver and thick restart),
``OOC4_alpha_beta_betak_k_TrdEigvalSel_TrdEigvecSel`` :320-351 (arrowhead Rayleigh-Ritz,
smallest k unfiltered / largest k filtered), ``OOC7_A_RitzV_u_alpha_beta_istart_iend_VOID``
:415-522 (Lanczos steps: alpha = v.u, CGS pass, beta = ||u|| taken before the second pass,
CGS pass 2, v_next = u / beta).

Device side: the basis V [ncv, n] float64 is row-major (one Lanczos vector per row); the
restart combination V[:k] = W^T V is ``inch_rr_combine``. Each step orthogonalises with
delayed CGS2 (``inch_dcgs2_step``, the default): the reference's second CGS pass of step
i - 1 is done inside step i, in the same read of V as step i's first pass, so a step reads V
twice instead of three times; the last residual of a cycle gets its second pass before the
restart. ``dcgs2=False`` selects the fused CGS2 kernels instead (the reference's order of
passes). Reductions are deterministic. The (ncv x ncv) projected eigenproblem of a restart
is solved on the device (torch's batched eigh, padded to the order where it is fast); the
in-cycle estimates, computed while the next segment runs, use the host form (select_ritz).
"""
import time

import numpy as np
import torch

from . import ops


def select_ritz(alpha, beta, beta_k, k, filtered):
    t = np.diag(alpha) + np.diag(beta[:-1], 1) + np.diag(beta[:-1], -1)
    if beta_k is not None:
        t[k, :k] = beta_k
        t[:k, k] = beta_k
    S, W = np.linalg.eigh(t)
    idx = np.argsort(S)
    sel = idx[-k:][::-1] if filtered else idx[:k]
    return S[sel], W[:, sel]


EIGH_MIN_ORDER = 72   # the device's batched eigh is about 20x faster from order 65 on (measured:
                      # 512 matrices of order 64 take 209 ms, of order 65 11 ms, of 72 4.6 ms)


def _projected_ritz_device(alpha, beta, m, Sel_prev, bk, k, filtered):
    """The projected eigenproblems of a restart on the device, for one structure or a batch
    (S structures; the stream is idle at a restart: a batched eigh of 60 projected matrices of
    order 84 takes 2 ms on the GPU against 23 ms on the host). The tridiagonal alpha / beta,
    with the thick-restart arrowhead, as in select_ritz; the largest k pairs when filtered,
    else the smallest k. Matrices of order below EIGH_MIN_ORDER are padded to it with a
    diagonal block of entries beyond the Gershgorin bound on the unwanted side (below the
    spectrum when the largest pairs are wanted, above it otherwise), whose eigenpairs are
    never selected and whose rows are dropped. alpha / beta: device [>= m, S]; Sel_prev
    [S, k] (host) replaces the first k diagonal entries and zeroes the first k betas after a
    thick restart, bk [S, k] (host) the arrowhead row. Returns (Sel [S, k] host, W [S, m, k]
    device contiguous, estimates [S, k] host)."""
    a = alpha[:m].T.contiguous()                   # [S, m]
    b = beta[:m].T.contiguous()
    dev = a.device
    if Sel_prev is not None:
        a[:, :k] = torch.as_tensor(np.ascontiguousarray(Sel_prev), device=dev)
        b[:, :k] = 0.0
    mp = max(m, EIGH_MIN_ORDER)
    S = a.shape[0]
    T = torch.zeros((S, mp, mp), dtype=a.dtype, device=dev)
    idx = torch.arange(m, device=dev)
    T[:, idx, idx] = a
    T[:, idx[:-1], idx[1:]] = b[:, :-1]
    T[:, idx[1:], idx[:-1]] = b[:, :-1]
    if bk is not None:
        bkd = torch.as_tensor(np.ascontiguousarray(bk), device=dev)
        T[:, k, :k] = bkd
        T[:, :k, k] = bkd
    if mp > m:
        g = T[:, :m, :m].abs().sum(-1).amax(-1)      # Gershgorin: every eigenvalue lies in [-g, g]
        pad = torch.arange(1, mp - m + 1, dtype=a.dtype, device=dev)
        fill = (g[:, None] + pad[None, :]) * (-1.0 if filtered else 1.0)
        ip = torch.arange(m, mp, device=dev)
        T[:, ip, ip] = fill
    w, Z = torch.linalg.eigh(T)                    # ascending per structure
    if filtered:
        Sel, W = w.flip(-1)[:, :k], Z.flip(-1)[:, :m, :k]
    else:
        Sel, W = w[:, :k], Z[:, :m, :k]
    W = W.contiguous()
    est = b[:, -1:] * W[:, -1, :]
    return Sel.cpu().numpy(), W, est.cpu().numpy()


class TRLMResult:
    def __init__(self, eigval, eigvec, iters, res, n_spmv, seconds, history):
        self.eigval = eigval
        self.eigvec = eigvec
        self.iters = iters
        self.res = res
        self.n_spmv = n_spmv
        self.seconds = seconds
        self.history = history


CYCLE_SEGMENT = 16   # Lanczos steps per captured segment of a restart cycle


class TRLM:
    """Device TRLM on an ``Operator`` (see operator.py).

    ``use_graph``: a restart cycle (restart vector, first extension step and the Lanczos
    sweep k+1..ncv-1) launches the same kernels on the same buffers every restart, so it is
    captured once into HIP graphs, one per segment of CYCLE_SEGMENT steps, and replayed;
    only beta_k is written from the host before a cycle.

    A cycle may end after any segment (the thick restart of the reference, from the m basis
    vectors built so far): with ``stop_est`` set, the projected matrix after each segment
    gives the Lanczos residual estimates |beta_m s_m,i| of the wanted Ritz pairs. The
    estimate e_s of segment s is computed on the host from an asynchronous copy of alpha /
    beta while segment s + 1 runs; the cycle ends after segment s + 1 once e_s <= stop_est
    or the geometric extrapolation e_s^2 / e_(s-1) to segment s + 1 is <= 0.3 stop_est."""

    def __init__(self, op, k, factor=2, ncv=None, use_graph=True, dcgs2=True):
        n = op.n
        if k >= n:
            raise AssertionError("ABORTED. k must be smaller than n")
        self.op = op
        self.k = int(k)
        self.n = n
        self.ncv = int(ncv) if ncv else min(max(int(round(factor * self.k)), self.k + 32), n - 1)
        dev = op.A.device
        self.V = torch.empty((self.ncv, n), dtype=torch.float64, device=dev)
        self.Vk = torch.empty((self.k, n), dtype=torch.float64, device=dev)
        self.u = torch.empty(n, dtype=torch.float64, device=dev)
        self.h = torch.empty(self.ncv + 1, dtype=torch.float64, device=dev)
        self.alpha = torch.zeros(self.ncv, dtype=torch.float64, device=dev)
        self.beta = torch.zeros(self.ncv, dtype=torch.float64, device=dev)
        self.scal = torch.empty(1, dtype=torch.float64, device=dev)
        self.betak_d = torch.empty(self.k, dtype=torch.float64, device=dev)
        self.ws = ops.Workspace(self.ncv + 1, dev)
        self.cgs_ws = ops.CGS2Workspace(min(self.ncv, ops.CGS2_MAX_J), dev)
        # delayed reorthogonalisation: two reads of V per step (the CGS2 forms take three)
        self.dcgs2 = bool(dcgs2)
        if self.dcgs2:
            self.dcgs_ws = ops.DCGS2Workspace(self.ncv, dev)
        self.use_graph = bool(use_graph) and op.timer is None
        self.graphs = {}
        self.stop_est = None    # residual-estimate threshold that ends a cycle early
        self.steps = 0          # Lanczos steps taken so far
        self.last_est = None    # max |beta_k| of the last restart
        self._ah = torch.empty(self.ncv, dtype=torch.float64, pin_memory=True)
        self._bh = torch.empty(self.ncv, dtype=torch.float64, pin_memory=True)

    # -- Lanczos steps i_start .. i_end-1: every step leaves its residual in u and, while the
    # basis has room (i + 1 < ncv), the next basis vector u / beta_i in V[i + 1]. With DCGS2
    # that vector is provisional until the next step corrects it (see _finish_cycle).
    def lanczos(self, i_start, i_end):
        V, u, h, ws = self.V, self.u, self.h, self.ws
        for i in range(i_start, i_end):
            self.op.apply(V[i], u)
            if self.dcgs2:
                last = i >= self.ncv - 1
                ops.dcgs2_step(V, i + 1, u, self.alpha, self.beta, self.beta[i:i + 1],
                               vnext=None if last else V[i + 1], ws=self.dcgs_ws)
                continue
            ops.multidot(V, i + 1, u, h, ws)
            self.alpha[i:i + 1].copy_(h[i:i + 1])
            last = i >= self.ncv - 1
            if i + 1 <= ops.CGS2_MAX_J:
                # fused: update + second-pass dots in one read of V, then update + normalise
                ops.cgs2_reorth(V, i + 1, h, u, self.beta[i:i + 1], vnext=None if last else V[i + 1],
                                ws=self.cgs_ws)
                continue
            ops.gemv_update(V, i + 1, h, u, 1.0, nrm2_out=self.beta[i:i + 1], ws=ws)
            ops.multidot(V, i + 1, u, h, ws)
            ops.gemv_update(V, i + 1, h, u, 1.0)
            if last:
                break
            ops.scale_div(u, self.beta[i:i + 1], V[i + 1])

    def _segments(self):
        """Step ranges of a restart cycle: k+1 .. ncv-1 in pieces of CYCLE_SEGMENT steps."""
        b = list(range(self.k + 1, self.ncv, CYCLE_SEGMENT)) + [self.ncv]
        return list(zip(b[:-1], b[1:]))

    def _segment(self, s, lo, hi):
        """Segment s of a restart cycle; the first one starts with the restart vector
        (orthogonalised against the k Ritz vectors) and the arrowhead extension step."""
        if s == 0:
            k = self.k
            V, u, ws = self.V, self.u, self.ws
            ops.multidot(V, k, u, self.h, ws)
            ops.gemv_update(V, k, self.h, u, 1.0, nrm2_out=self.scal, ws=ws)
            ops.scale_div(u, self.scal, V[k])
            self.op.apply(V[k], u)
            ops.multidot(V[k:k + 1], 1, u, self.alpha[k:k + 1], ws)
            ops.axpy_dev(self.alpha[k:k + 1], -1.0, V[k], u)
            ops.gemv_update(V, k, self.betak_d, u, 1.0, nrm2_out=self.beta[k:k + 1], ws=ws)
            ops.scale_div(u, self.beta[k:k + 1], V[k + 1])
        self.lanczos(lo, hi)

    def _run_segment(self, s, lo, hi):
        if not self.use_graph:
            self._segment(s, lo, hi)
            return
        if s not in self.graphs:
            n0, a0 = self.op.n_spmv, self.op.n_apply
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._segment(s, lo, hi)
            self.graphs[s] = (g, self.op.n_spmv - n0, self.op.n_apply - a0)
        else:
            self.op.n_spmv += self.graphs[s][1]
            self.op.n_apply += self.graphs[s][2]
        self.graphs[s][0].replay()

    def _run_cycle(self, S, beta_k, filtered):
        """One restart cycle; returns the number m of basis vectors built (alpha[:m],
        beta[:m] hold the projected matrix, u the residual)."""
        k = self.k
        segs = self._segments()
        pending = None   # (m, event) of the last segment whose alpha / beta copy is in flight
        prev = None      # estimate of the segment before it
        for s, (lo, hi) in enumerate(segs):
            self._run_segment(s, lo, hi)
            if hi >= self.ncv:
                return hi
            if self.stop_est is None:
                continue
            if pending is not None:
                m, ev = pending
                ev.synchronize()
                a = self._ah[:m].numpy().copy()
                b = self._bh[:m].numpy().copy()
                a[:k] = S
                b[:k] = 0.0
                _, W = select_ritz(a, b, beta_k, k, filtered)
                est = float(np.max(np.abs(b[-1] * W[-1, :])))
                if est <= self.stop_est or (prev is not None and est < prev
                                            and est * est / prev <= 0.3 * self.stop_est):
                    return hi
                prev = est
            self._ah[:hi].copy_(self.alpha[:hi], non_blocking=True)
            self._bh[:hi].copy_(self.beta[:hi], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            pending = (hi, ev)
        return self.ncv

    def _finish_cycle(self, m):
        """With DCGS2 the last residual u has had one orthogonalisation pass (its second one
        would come in the next step): give it that pass against V[:m] before the restart."""
        if self.dcgs2:
            ops.multidot(self.V, m, self.u, self.h, self.ws)
            ops.gemv_update(self.V, m, self.h, self.u, 1.0)

    def _restart_rr(self, m, S_prev, beta_k, filtered):
        """Thick restart from V[:m]: the projected eigenproblem (alpha / beta on the device,
        after a restart the previous Ritz values S_prev and the arrowhead beta_k) solved on the
        device, the k Ritz vectors formed in V[:k]."""
        k = self.k
        Sel, Wd, est = _projected_ritz_device(self.alpha.view(-1, 1), self.beta.view(-1, 1), m,
                                              None if S_prev is None else np.asarray(S_prev)[None, :],
                                              None if beta_k is None else np.asarray(beta_k)[None, :], k, filtered)
        ops.rr_combine(self.V, m, Wd[0], self.Vk)
        self.V[:k].copy_(self.Vk)
        S, beta_k = Sel[0], est[0]
        self.last_est = float(np.max(np.abs(beta_k)))
        return S, beta_k, float(np.linalg.norm(beta_k))

    def run(self, maxiter=10000, tol=0.0, v0=None, seed=0, filtered=False, check=None, verbose=False,
            stop_est=None):
        """Returns TRLMResult. ``check(self, S)`` (optional) is called after every restart and
        may return True to stop (residual-verified acceptance; it may also update
        ``self.stop_est``); ``tol`` is the reference's absolute criterion on ||beta_k||;
        ``stop_est`` (optional) ends restart cycles early, see the class docstring."""
        self.stop_est = stop_est
        t0 = time.time()
        k, ncv = self.k, self.ncv
        if tol == 0:
            tol = np.finfo(np.float64).eps
        if v0 is None:
            g = torch.Generator(device=self.V.device)
            g.manual_seed(int(seed))
            self.u.copy_(torch.rand(self.n, generator=g, dtype=torch.float64, device=self.V.device))
        else:
            self.u.copy_(torch.as_tensor(v0, dtype=torch.float64, device=self.V.device).reshape(-1))
        self.V[0].copy_(self.u / torch.linalg.norm(self.u))
        self.lanczos(0, ncv)
        self._finish_cycle(ncv)
        self.steps = ncv
        S, beta_k, res = self._restart_rr(ncv, None, None, filtered)
        history = [res]
        it = 0
        stop = False
        self.last_res = res
        if check is not None:
            stop = bool(check(self, S))
        for it in range(maxiter):
            if res <= tol or stop:
                break
            self.betak_d.copy_(torch.as_tensor(beta_k, device=self.V.device))
            m = self._run_cycle(S, beta_k, filtered)
            self._finish_cycle(m)
            self.steps += m - k
            S, beta_k, res = self._restart_rr(m, S, beta_k, filtered)
            history.append(res)
            self.last_res = res
            if verbose:
                print("TRLM restart %d  ||beta_k|| = %.3e  spmv = %d" % (it, res, self.op.n_spmv))
            if check is not None:
                stop = bool(check(self, S))
        S = np.asarray(S, dtype=np.float64)
        torch.cuda.synchronize()
        return TRLMResult(S, self.V[:k], it, res, self.op.n_spmv, time.time() - t0, history)


def rayleigh_quotients(A, Vrows, precision="f64"):
    """vi . (A vi) for the rows of Vrows (float64 [k, n]) with the unfiltered operator."""
    X = Vrows.T.contiguous()
    Y = A.matmat(X, precision=precision) if precision else A.matfree(X)
    return (X * Y).sum(0)
