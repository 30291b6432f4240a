"""Batched solve of many independent structures in one launch sequence.

The reference solves one structure per iteration of its driver loop
(``Command8A/commandBenchmark_InchingTRLMHD_Linux_Pdb_0064_A100.py:189``). Here all
structures of a batch form one block-diagonal ANM matrix: their coordinates are placed
far apart in one frame, so a single cell-list search, RCM ordering (the BFS keeps every
connected structure contiguous) and assembly build all of them. Thick-restart Lanczos
then runs every structure in lockstep: the operator is one launch over the whole batch,
and every inner product, norm, update and restart combination is a *segmented* kernel
over the structures' dof ranges (``inch_seg_*``; each Lanczos step orthogonalises with the
segmented delayed CGS2, two reads of the basis). A restart cycle of all structures is
captured once into HIP graphs, one per segment of steps, and replayed. The projected
eigenproblems (one arrowhead matrix per structure) are solved on the host as one batched
``numpy.linalg.eigh``; between segments they give the Lanczos residual estimates that end
a cycle as soon as every unconverged structure is predicted to meet its target.

Each structure's latency is recorded when it meets the tolerance; the batch runs until
all have. The rigid-body block of every structure is deflated with a Hotelling shift, as
in the reference's TRLMHD.
"""
import time

import numpy as np
import torch

from . import ops
from ._lib import call
from .hessian import build_hessian
from .operator import Operator
from .trlm import EIGH_MIN_ORDER, _projected_ritz_device  # noqa: F401  (re-exported)


class Segments:
    """Per-structure dof ranges cut into chunks of at most ``chunk`` dofs."""

    def __init__(self, seg_ptr, device, chunk=4096):
        seg_ptr = np.asarray(seg_ptr, dtype=np.int64)
        self.S = seg_ptr.shape[0] - 1
        cb, ce, cs, sp = [], [], [], [0]
        for s in range(self.S):
            b, e = int(seg_ptr[s]), int(seg_ptr[s + 1])
            for x in range(b, e, chunk):
                cb.append(x)
                ce.append(min(e, x + chunk))
                cs.append(s)
            sp.append(len(cb))
        self.nchunk = len(cb)
        self.seg_ptr = seg_ptr
        self.cbeg = torch.as_tensor(np.asarray(cb, dtype=np.int64), device=device)
        self.cend = torch.as_tensor(np.asarray(ce, dtype=np.int64), device=device)
        self.cseg = torch.as_tensor(np.asarray(cs, dtype=np.int32), device=device)
        self.sptr = torch.as_tensor(np.asarray(sp, dtype=np.int64), device=device)
        self.sizes = np.diff(seg_ptr)

    def work(self, J, device):
        return torch.empty(self.nchunk * max(1, J), dtype=torch.float64, device=device)


P = ops.P
SP = ops.stream_ptr


class SegOps:
    def __init__(self, seg, maxJ, device):
        self.seg = seg
        self.maxJ = int(maxJ)
        self.wk = seg.work(maxJ, device)
        self.dJ = min(self.maxJ, 256)
        nw = call("inch_seg_dcgs2_work_size", seg.nchunk, seg.S, self.dJ)
        self.wk2 = torch.empty(int(nw), dtype=torch.float64, device=device)

    def dcgs2(self, V, J, u, alpha_i, beta_prev, beta_out):
        """Delayed CGS2 step per segment (inch_seg_dcgs2_step): corrects the provisional V[J-1]
        against V[:J-1], projects u = Op V[J-1] off the corrected V[:J]; writes alpha_i[s]
        (using beta_prev[s]) and beta_out[s] = |u_s| (two reads of V)."""
        s = self.seg
        if J > self.dJ:
            raise ValueError("dcgs2: workspace holds J <= %d" % self.dJ)
        call("inch_seg_dcgs2_step", s.nchunk, s.S, P(s.cbeg), P(s.cend), P(s.cseg), P(s.sptr), int(J), P(V),
             V.stride(0), P(u), P(alpha_i), P(beta_prev), P(beta_out), P(self.wk2), SP())

    def cgs_b(self, V, J, h, hld, u, h2, nrm):
        """u -= V[:J] h per segment; h2[s] = (V[:J] u, |u|^2) over segment s; nrm[s] = |u| (one
        read of V)."""
        s = self.seg
        if J + 1 > self.maxJ:
            raise ValueError("cgs_b: workspace holds J + 1 <= %d partials per chunk" % self.maxJ)
        call("inch_seg_cgs_b", s.nchunk, s.S, P(s.cbeg), P(s.cend), P(s.cseg), P(s.sptr), int(J), P(V), V.stride(0),
             P(h), int(hld), P(u), P(h2), P(nrm), P(self.wk), SP())

    def multidot(self, V, J, u, h):
        s = self.seg
        call("inch_seg_multidot", s.nchunk, s.S, P(s.cbeg), P(s.cend), P(s.sptr), int(J), P(V), V.stride(0), P(u),
             P(h), P(self.wk), SP())

    def update(self, V, J, h, hld, scale, u, nrm=None):
        s = self.seg
        call("inch_seg_update", s.nchunk, s.S, P(s.cbeg), P(s.cend), P(s.cseg), P(s.sptr), int(J),
             P(V) if J > 0 else None, V.stride(0) if J > 0 else u.numel(), P(h) if J > 0 else None, int(hld),
             float(scale), P(u), P(nrm), P(self.wk) if nrm is not None else None, SP())

    def rigid(self, rig, x, y, scale):
        """y -= scale * P x, P the per-segment rigid-body projection of ``rig`` (a RigidBody of
        the same segments); x may be y. With y None only the moments of x are formed, into
        rig.g [S, 6] (sum of x's components, sum r_i x x_i with r_i about rig.cen)."""
        s = self.seg
        call("inch_seg_rigid_apply", s.nchunk, s.S, P(s.cbeg), P(s.cend), P(s.cseg), P(s.sptr), P(rig.xyz),
             P(rig.cen), P(rig.kinv), P(rig.ninv), P(x), P(y), float(scale), P(rig.g), P(self.wk), SP())

    def scale(self, src, div, dst, dst2=None):
        s = self.seg
        call("inch_seg_scale", s.nchunk, P(s.cbeg), P(s.cend), P(s.cseg), P(src), P(div), P(dst), P(dst2), SP())

    def combine(self, V, m, W, out):
        s = self.seg
        k = W.shape[2]
        call("inch_seg_combine", s.nchunk, P(s.cbeg), P(s.cend), P(s.cseg), int(m), int(k), P(V), V.stride(0),
             P(W), P(out), out.stride(0), SP())

    def resnorm(self, X, Y, theta, out):
        s = self.seg
        k = X.shape[0]
        call("inch_seg_resnorm", s.nchunk, s.S, P(s.cbeg), P(s.cend), P(s.cseg), P(s.sptr), int(k), P(X),
             X.stride(0), P(Y), Y.stride(0), P(theta), P(out), P(self.wk), SP())


def place_apart(structures, gap=30.0):
    """Translate each structure (centred) onto a cubic grid with cells wider than every
    structure plus ``gap`` A, so no pair of atoms from different structures is within the
    cutoff. Returns (xyz [sum N, 3], atom offsets [S+1])."""
    S = len(structures)
    span = max(float(np.max(np.ptp(x, axis=0))) for x in structures) + gap
    g = int(np.ceil(S ** (1.0 / 3.0)))
    parts, off = [], [0]
    for s, x in enumerate(structures):
        c = np.array([s % g, (s // g) % g, s // (g * g)], dtype=np.float64) * span
        parts.append(x - x.mean(0) + c)
        off.append(off[-1] + x.shape[0])
    return np.concatenate(parts), np.asarray(off, dtype=np.int64)


def segmented_rigid_basis(xyz, seg_atoms, segops):
    """6 rigid-body rows per structure, nonzero only on its own dofs, orthonormalised per
    structure (two Cholesky-QR sweeps with segmented Gram products)."""
    n = xyz.shape[0]
    dev = xyz.device
    S = len(seg_atoms) - 1
    comp = torch.as_tensor(np.repeat(np.arange(S), np.diff(seg_atoms)), device=dev)
    cnt = torch.as_tensor(np.diff(seg_atoms).astype(np.float64), device=dev)
    Q = torch.zeros((6, 3 * n), dtype=torch.float64, device=dev)
    inv = (1.0 / torch.sqrt(cnt))[comp]
    for a in range(3):
        Q[a].view(n, 3)[:, a] = inv
    x = xyz
    zero = torch.zeros_like(x[:, 0])
    Q[3] = torch.stack([zero, -x[:, 2], x[:, 1]], 1).reshape(-1)
    Q[4] = torch.stack([x[:, 2], zero, -x[:, 0]], 1).reshape(-1)
    Q[5] = torch.stack([-x[:, 1], x[:, 0], zero], 1).reshape(-1)
    G = torch.empty((6, S * 6), dtype=torch.float64, device=dev)
    for _ in range(2):
        for a in range(6):
            segops.multidot(Q, 6, Q[a].contiguous(), G[a])
        Gh = G.cpu().numpy().reshape(6, S, 6).transpose(1, 0, 2)    # [S, 6, 6]
        Gh = 0.5 * (Gh + Gh.transpose(0, 2, 1))
        R = np.linalg.cholesky(Gh).transpose(0, 2, 1)                 # G = R^T R
        W = torch.as_tensor(np.ascontiguousarray(np.linalg.inv(R)), device=dev)   # [S, 6, 6]
        Qn = torch.empty_like(Q)
        segops.combine(Q, 6, W, Qn)
        Q = Qn
    return Q


class RigidBody:
    """Per-segment data of the rigid-body projection computed from the coordinates
    (``SegOps.rigid``): centroids, inverse rotational Gram matrices K_s^{-1} =
    (sum_i |r_i|^2 I - r_i r_i^T)^{-1} and inverse atom counts. The projection equals
    x -> Q^T Q x for the orthonormal rows of ``segmented_rigid_basis`` and reads the
    coordinates (24 bytes per atom) instead of six basis rows (144 bytes per atom).

    The centroids and K_s come from the same segmented moment kernel (no atomics: a scatter-add
    of every atom into its segment's nine entries took seconds for 512 segments): the moments
    of the coordinates about the origin give the centroids, and those of the rotation field
    e_l x r_i give K_s e_l = sum_i r_i x (e_l x r_i)."""

    def __init__(self, xyz, seg_atoms, so):
        dev = xyz.device
        seg_atoms = np.asarray(seg_atoms, dtype=np.int64)
        S = len(seg_atoms) - 1
        cnt = np.diff(seg_atoms)
        if np.any(cnt < 3):
            raise ValueError("every segment needs at least 3 atoms for its rigid-body basis")
        if so.seg.S != S or so.seg.nchunk == 0 or int(so.seg.seg_ptr[-1]) != 3 * int(seg_atoms[-1]):
            raise ValueError("the segment operations must cover the atoms' dofs")
        self.xyz = xyz.contiguous()
        self.g = torch.empty(S * 6, dtype=torch.float64, device=dev)
        self.kinv = self.ninv = None
        self.cen = torch.zeros((S, 3), dtype=torch.float64, device=dev)
        so.rigid(self, self.xyz.reshape(-1), None, 0.0)            # moments about the origin
        cnt_d = torch.as_tensor(cnt.astype(np.float64), device=dev)
        self.cen = (self.g.view(S, 6)[:, :3] / cnt_d[:, None]).contiguous()
        comp = torch.repeat_interleave(torch.arange(S, device=dev), torch.as_tensor(cnt, device=dev))
        r = self.xyz - self.cen[comp]
        zero = torch.zeros_like(r[:, 0])
        K = torch.empty((S, 3, 3), dtype=torch.float64, device=dev)
        for l, field in enumerate(((zero, -r[:, 2], r[:, 1]), (r[:, 2], zero, -r[:, 0]), (-r[:, 1], r[:, 0], zero))):
            so.rigid(self, torch.stack(field, 1).reshape(-1), None, 0.0)
            K[:, :, l] = self.g.view(S, 6)[:, 3:]
        # the S 3 x 3 determinants and inverses on the host (numpy; S x 9 numbers)
        Kh = K.cpu().numpy()
        Kh = 0.5 * (Kh + Kh.transpose(0, 2, 1))
        if np.any(np.abs(np.linalg.det(Kh)) <= 1e-12 * np.linalg.norm(Kh, axis=(1, 2)) ** 3):
            raise ValueError("a segment's atoms are collinear: its rotations do not span three directions")
        self.kinv = torch.as_tensor(np.linalg.inv(Kh).reshape(S, 9), device=dev)
        self.ninv = torch.as_tensor(1.0 / cnt.astype(np.float64), device=dev)


class BatchResult:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def solve_batch(structures, k=20, rc=8.0, plus_i=1.0, precision="mf64", tol_rel=1e-8, shift=10.0, maxiter=500,
                factor=2, seed=0, use_graph=True, verbose=False, pol=None, dcgs2=True):
    """Lowest k non-trivial modes of every structure (list of [N_s, 3] coordinate arrays, A).

    Returns BatchResult with per-structure eigenvalues [S, k] (PlusI included), eigenvectors
    (rows [k, dof] in the batch's RCM order), relative residuals [S, k], per-structure
    latency (seconds from start until its convergence) and the total time."""
    ops.require_gpu()
    dev = torch.device("cuda", torch.cuda.current_device())
    torch.cuda.synchronize()
    t0 = time.time()
    X, off = place_apart(structures)
    prec = {"mf64": (), "f32": ("f32",), "f64": ("f64",), "mixed": ("f32",)}[precision]
    A = build_hessian(X, rc, plus_i=plus_i, reorder=True, precision=prec, device=dev)
    # structures stay contiguous after RCM; locate each by its first atom's new position
    src_of = np.repeat(np.arange(len(structures)), np.diff(off))[A.order]       # structure of each new row
    starts = np.nonzero(np.r_[True, src_of[1:] != src_of[:-1]])[0]
    seg_struct = src_of[starts]
    seg_atoms = np.r_[starts, A.n_atoms].astype(np.int64)
    if len(starts) != len(structures):
        raise RuntimeError("structures are not contiguous after ordering (are they closer than the cutoff?)")
    return solve_segments(A, seg_atoms, k, seg_struct=seg_struct, precision=precision, tol_rel=tol_rel, shift=shift,
                          maxiter=maxiter, factor=factor, seed=seed, use_graph=use_graph, verbose=verbose, t0=t0,
                          pol=pol, dcgs2=dcgs2)


def components(A):
    """Connected components of the neighbour graph of A as contiguous atom ranges of the
    RCM order: the component start offsets [C+1]. The device ordering visits each component
    completely before the next one and reports where each starts."""
    comps = getattr(A, "components", None)
    if comps is None:
        raise ValueError("components need a matrix built with reorder=True and the device ordering")
    return np.asarray(comps, dtype=np.int64)


CYCLE_SEGMENT = 16   # lockstep Lanczos steps per captured segment of a cycle (8 with a filter,
                     # whose steps cost a filter application each)
STOP_SAFETY = 0.3    # a filtered cycle ends when the calibrated estimate predicts 0.3 x the target
STALL = 0.7          # a float32 phase ends when its worst residual improves by less than this factor


def _projected_ritz(a, b, bk, k, filtered):
    """Ritz values, eigenvector columns and residual estimates of every structure's projected
    matrix: tridiagonal alpha / beta ([m, S]), with the arrowhead row beta_k [S, k] after a
    thick restart (``TRLM/T1.py:320-351``). Largest k for a filtered operator, else smallest k.
    Returns (Sel [S, k], W [S, m, k], estimates beta_{m-1} W[:, m-1, :] [S, k])."""
    m, S = a.shape
    T = np.zeros((S, m, m))
    idx = np.arange(m)
    T[:, idx, idx] = a.T
    T[:, idx[:-1], idx[1:]] = b[:-1].T
    T[:, idx[1:], idx[:-1]] = b[:-1].T
    if bk is not None:
        T[:, k, :k] = bk
        T[:, :k, k] = bk
    w, Z = np.linalg.eigh(T)                 # ascending per structure
    if filtered:
        Sel, W = w[:, ::-1][:, :k], Z[:, :, ::-1][:, :, :k]
    else:
        Sel, W = w[:, :k], Z[:, :, :k]
    return Sel, W, b[-1][:, None] * W[:, -1, :]


def solve_segments(A, seg_atoms, k, seg_struct=None, precision="mf64", tol_rel=1e-8, shift=10.0, maxiter=500,
                   factor=2, seed=0, use_graph=True, verbose=False, t0=None, pol=None, dcgs2=True,
                   segment=None, early_stop=True, phase1_rel=3e-7, v0=None, timer=None, band=None,
                   deflate=None):
    """Lockstep thick-restart Lanczos over the contiguous atom segments of a block-diagonal
    A: k modes per segment, per-segment rigid-body Hotelling deflation, optional common
    Chebyshev filter ``pol`` (largest k of p(A) per segment, then Rayleigh quotients).

    A restart cycle runs as HIP-graph segments of ``segment`` steps. With ``early_stop`` the
    host forms every structure's projected matrix from an asynchronous copy of alpha / beta
    after each segment (while the next one runs) and ends the cycle (a thick restart from the
    m vectors built) once every unconverged structure's largest Lanczos residual estimate
    |beta_m s_m,i|, or its geometric extrapolation one segment ahead, meets the threshold:
    tol_rel * lam_max without a filter (the estimate is the residual of A), and for a filtered
    operator STOP_SAFETY * tol_rel scaled by the ratio of estimate to float64 residual that the
    previous cycle's float64 check measured per structure (1 before any check).

    ``precision``: 'mf64' / 'f64' / 'f32' operators throughout, or 'mixed': the float32 block
    operator until every structure's float64-checked residual reaches ``phase1_rel`` or stops
    improving, then float64 (matrix-free) filtered subspace steps on the Ritz block with a
    float64 Rayleigh-Ritz per structure until ``tol_rel``. ``v0``: optional float64 start
    vector [n] (e.g. a combination of predicted modes); random otherwise. ``timer``: list
    that receives HIP event pairs around the operator launches (graphs are then off).

    ``deflate``: 'shift' adds -shift Q Q^T to the operator (the reference's Hotelling
    deflation); 'project' applies (I - Q Q^T) after the operator instead, so the rigid-body
    directions become zero eigenvalues of the operator rather than outliers at -shift (a
    filtered operator's rigid modes otherwise sit far below its spectrum, which costs Lanczos
    steps). None (default): 'project' with a filter (whose wanted end is the top, far from
    zero), 'shift' without one (the wanted end is the bottom, where zero would be).

    ``band`` = (p_r, r) for a filtered operator: only the pairs inside the filter's band
    [PlusI, r] have to converge -- during a cycle those whose Ritz value of p(A) is at least
    p_r = p(r), after a float64 check those whose Rayleigh quotient is at most r; the others
    are returned unconverged. ``in_band`` in the result marks the pairs inside."""
    dev = A.device
    if t0 is None:
        torch.cuda.synchronize()
        t0 = time.time()
    if seg_struct is None:
        seg_struct = np.arange(len(seg_atoms) - 1)
    if precision not in ("mf64", "f64", "f32", "mixed"):
        raise ValueError("precision must be 'mf64', 'f64', 'f32' or 'mixed'")
    filtered = pol is not None
    if segment is None:
        segment = CYCLE_SEGMENT // 2 if filtered else CYCLE_SEGMENT
    if filtered and shift > 0:
        shift = -40.0
    S = len(seg_atoms) - 1
    n = A.shape[0]
    seg = Segments(3 * seg_atoms, dev)
    min_dof = int(seg.sizes.min())
    ncv = min(max(int(round(factor * k)), k + 32), min_dof - 7)
    if ncv <= k + 1:
        raise AssertionError("ABORTED. k must be smaller than every structure's dof count minus the rigid modes")
    so = SegOps(seg, max(ncv, 6) + 1, dev)
    dcgs2 = bool(dcgs2) and ncv <= so.dJ
    rig = RigidBody(A.xyz, seg_atoms, so)
    mixed = precision == "mixed"
    op = Operator(A, "f32" if mixed else precision, pol=pol, timer=timer)
    if timer is not None:
        use_graph = False
    f64 = dict(dtype=torch.float64, device=dev)
    V = torch.empty((ncv, n), **f64)
    Vk = torch.empty((k, n), **f64)
    u = torch.empty(n, **f64)
    h = torch.empty(S * (ncv + 1), **f64)
    h2 = torch.empty(S * (ncv + 1), **f64)
    alpha = torch.zeros((ncv, S), **f64)
    beta = torch.zeros((ncv, S), **f64)
    scal = torch.empty(S, **f64)
    betak_d = torch.zeros((S, k), **f64)
    ah = torch.empty((ncv, S), dtype=torch.float64, pin_memory=True)
    bh = torch.empty((ncv, S), dtype=torch.float64, pin_memory=True)

    if deflate is None:
        deflate = "project" if filtered else "shift"
    project = deflate == "project"
    if deflate not in ("shift", "project"):
        raise ValueError("deflate must be 'shift' or 'project'")

    def apply(x, y):
        op.apply(x, y)
        if project:
            so.rigid(rig, y, y, 1.0)
        elif shift:
            so.rigid(rig, x, y, -shift)

    def lanczos(i0, i1):
        for i in range(i0, i1):
            apply(V[i], u)
            J = i + 1
            if dcgs2:
                # delayed CGS2: V[i] is corrected here, its successor at the next step
                so.dcgs2(V, J, u, alpha[i], beta[i - 1] if i > 0 else None, beta[i])
            else:
                so.multidot(V, J, u, h)
                alpha[i].copy_(h[:S * J].view(S, J)[:, i])
                # CGS2: update + second projection in one read of V (beta from the first
                # update), then the second update
                so.cgs_b(V, J, h, J, u, h2, beta[i])
                so.update(V, J, h2, J + 1, 1.0, u)
            if i + 1 < ncv:
                so.scale(u, beta[i], V[i + 1])

    def head():
        # restart vector orthogonalised against the k Ritz vectors, then the arrowhead step
        so.multidot(V, k, u, h)
        so.update(V, k, h, k, 1.0, u, nrm=scal)
        so.scale(u, scal, V[k])
        apply(V[k], u)
        so.multidot(V[k:k + 1], 1, u, alpha[k])
        so.update(V[k:k + 1], 1, alpha[k], 1, 1.0, u)
        so.update(V, k, betak_d, k, 1.0, u, nrm=beta[k])
        so.scale(u, beta[k], V[k + 1])

    graphs = {}

    def run_segment(first, s, lo, hi):
        key = (s, op.precision)
        if not use_graph or first:          # the first sweep runs once: launched directly
            if not first and s == 0:
                head()
            lanczos(lo, hi)
            return
        if key not in graphs:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                if not first and s == 0:
                    head()
                lanczos(lo, hi)
            graphs[key] = g
        graphs[key].replay()

    def finish(m):
        # with delayed CGS2 the last residual has had one pass: give it the second
        if dcgs2:
            so.multidot(V, m, u, h)
            so.update(V, m, h, m, 1.0, u)

    steps = [0]

    def run_cycle(first, Sel, bk, thr, active):
        """One cycle (the first Lanczos sweep, or a thick-restart cycle from the k Ritz
        vectors in V[:k] and the residual in u); returns the number m of basis vectors built."""
        lo0 = 0 if first else k + 1
        b_ = list(range(lo0, ncv, segment)) + [ncv]
        pending, prev = None, None
        for s, (lo, hi) in enumerate(zip(b_[:-1], b_[1:])):
            run_segment(first, s, lo, hi)
            steps[0] += hi - lo + (0 if first or s else 1)
            if hi >= ncv:
                return ncv
            if thr is None:
                continue
            if pending is not None and pending[0] >= k + 2:
                m, ev = pending
                ev.synchronize()
                a, b = ah[:m].numpy().copy(), bh[:m].numpy().copy()
                if not first:
                    a[:k] = Sel.T
                    b[:k] = 0.0
                Sel_c, _, bk_c = _projected_ritz(a[:, active], b[:, active], None if first else bk[active], k,
                                                 filtered)
                est = band_max(np.abs(bk_c), Sel_c >= band[0]) if band is not None else np.abs(bk_c).max(1)
                th = thr[active]
                ok = est <= th
                if prev is not None:
                    ok |= (est < prev) & (est * est <= 0.3 * th * prev)
                if verbose:
                    print("  m=%d: worst estimate / threshold %.2e, %d of %d structures ready"
                          % (m, float(np.max(est / th)), int(ok.sum()), ok.size))
                if ok.all():
                    return hi
                prev = est
            ah[:hi].copy_(alpha[:hi], non_blocking=True)
            bh[:hi].copy_(beta[:hi], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            pending = (hi, ev)
        return ncv

    def band_max(vals, mask):
        """Per-structure maximum of vals over the masked pairs (0 where none)."""
        return np.where(mask, vals, 0.0).max(1)

    def restart(m, Sel, bk, first):
        Sel, Wd, bkn = _projected_ritz_device(alpha, beta, m, None if first else Sel, None if first else bk, k,
                                              filtered)
        so.combine(V, m, Wd, Vk)
        V[:k].copy_(Vk)
        return Sel, bkn

    def verify(X):
        """float64 check of the row block X [k, n]: per-structure eigenvalue estimates
        (Ritz values without a filter, Rayleigh quotients with one) and relative residuals."""
        Y = A.matfree(X.T.contiguous()).T.contiguous()
        rq = torch.empty((k, S), **f64)
        for q in range(k):
            so.multidot(X[q:q + 1], 1, Y[q], rq[q])
        theta_h = rq.T.cpu().numpy()                # [S, k]
        rn = torch.empty(S * k, **f64)
        so.resnorm(X, Y, torch.as_tensor(np.ascontiguousarray(theta_h), device=dev), rn)
        return theta_h, rn.cpu().numpy().reshape(S, k) / lam_max

    # lam_max for the relative residual: a Lanczos bound of the whole batch (the filter's
    # upper end when one is given)
    from .chebyshev import BOUND_STOP_REL, spectrum_bound
    lam_max = (spectrum_bound(A, precision="f32" if mixed else precision, stop_rel=BOUND_STOP_REL)[1]
               if pol is None else float(pol.cc + pol.dd))
    tol_abs = tol_rel * lam_max
    target = phase1_rel if mixed else tol_rel
    # in-cycle thresholds per structure, in the units of the Lanczos estimate
    thr = np.full(S, (target * lam_max) if not filtered else STOP_SAFETY * target)

    if v0 is None:
        g = torch.Generator(device=dev)
        g.manual_seed(int(seed))
        u.copy_(torch.rand(n, generator=g, dtype=torch.float64, device=dev))
    else:
        u.copy_(torch.as_tensor(v0, dtype=torch.float64, device=dev).reshape(-1))
    if project:
        so.rigid(rig, u, u, 1.0)
    so.update(V, 0, None, 1, 0.0, u, nrm=scal)
    so.scale(u, scal, V[0])
    done_at = np.full(S, np.nan)
    conv = np.zeros(S, dtype=bool)
    checked = None          # (theta [S, k], rel [S, k]) of the last float64 check of V[:k]
    hist = []
    Sel, bk = None, None
    cycles = 0              # restart cycles after the first sweep
    phase1_cycles = None
    for it in range(maxiter + 1):
        first = it == 0
        active = ~conv
        # in-cycle checks: always for a filtered operator; without a filter from the second
        # cycle on, once every unconverged structure's estimate is within reach (the host
        # eigh of every projected matrix per segment is not free for large batches)
        use_thr = early_stop and (filtered or (not first and bool(np.all(est[active] <= 1e4 * thr[active]))))
        if not first:
            betak_d.copy_(torch.as_tensor(bk, device=dev))
        m = run_cycle(first, Sel, bk, thr if use_thr else None, active)
        if not first:
            cycles += 1
        finish(m)
        Sel, bk = restart(m, Sel, bk, first)
        est = band_max(np.abs(bk), Sel >= band[0]) if band is not None else np.abs(bk).max(1)    # [S]
        if not filtered and not mixed:
            newly = active & (est <= tol_abs)
        else:
            checked = verify(V[:k])
            worst = band_max(checked[1], checked[0] <= band[1]) if band is not None else checked[1].max(1)
            newly = active & (worst <= target)
            # calibrate the in-cycle threshold: estimate per unit of float64 residual
            thr = np.where(worst > 0, STOP_SAFETY * target * est / np.maximum(worst, 1e-300), thr)
            hist.append(float(worst.max()))
        done_at[newly] = time.time() - t0
        conv |= newly
        if verbose:
            print("lockstep cycle %d (m=%d): %d/%d converged, max estimate %.2e%s"
                  % (it, m, conv.sum(), S, est.max(), "" if checked is None else ", worst float64 rel_res %.2e"
                     % checked[1].max()))
        if conv.all():
            break
        if mixed and len(hist) >= 2 and hist[-1] > STALL * hist[-2]:
            break                                   # the float32 operator's floor
    if mixed:
        phase1_cycles = cycles
        if filtered:
            checked = _fp64_block_refine(A, V, k, pol, rig, so, S, tol_rel, checked, verify, verbose=verbose,
                                         timer=timer, band=band)
        else:
            # no filter to refine with: a float64 lockstep run started from a combination of
            # every structure's float32 Ritz vectors
            gc = torch.Generator(device=dev)
            gc.manual_seed(int(seed) + 1)
            c = torch.rand(k, generator=gc, dtype=torch.float64, device=dev) + 0.5
            r2 = solve_segments(A, seg_atoms, k, seg_struct=seg_struct, precision="mf64", tol_rel=tol_rel,
                                shift=shift, maxiter=maxiter, factor=factor, seed=seed, use_graph=use_graph,
                                verbose=verbose, t0=t0, dcgs2=dcgs2, segment=segment, early_stop=early_stop,
                                v0=(c[:, None] * V[:k]).sum(0), timer=timer)
            r2.phase1_cycles = phase1_cycles
            r2.restarts += cycles
            r2.n_spmv += steps[0]
            r2.operator_launches += op.n_spmv
            return r2
        now = time.time() - t0
        conv = (band_max(checked[1], checked[0] <= band[1]) if band is not None else checked[1].max(1)) <= tol_rel
        done_at[:] = now
    # float64 check of every pair (matrix-free operator, segmented residual norms)
    Xr = V[:k]
    if checked is None:
        if filtered:
            checked = verify(Xr)
        else:
            # Ritz values of A itself; residuals with the matrix-free float64 operator
            Y = A.matfree(Xr.T.contiguous()).T.contiguous()
            rn = torch.empty(S * k, **f64)
            so.resnorm(Xr, Y, torch.as_tensor(np.ascontiguousarray(Sel), device=dev), rn)
            checked = (Sel, rn.cpu().numpy().reshape(S, k) / lam_max)
    theta_h, rel = checked
    torch.cuda.synchronize()
    total = time.time() - t0
    ordq = np.argsort(theta_h, axis=1, kind="stable")   # ascending eigenvalues per segment
    theta_h = np.take_along_axis(theta_h, ordq, axis=1)
    rel = np.take_along_axis(rel, ordq, axis=1)
    order = np.argsort(seg_struct)
    done = np.where(np.isnan(done_at), total, done_at)
    in_band = (theta_h <= band[1]) if band is not None else np.ones_like(theta_h, dtype=bool)
    return BatchResult(eigval=theta_h[order], rel_res=rel[order], latency=done[order], seconds=total,
                       restarts=cycles, n_spmv=steps[0], ncv=ncv, lam_max=lam_max, seg_atoms=seg_atoms,
                       seg_struct=seg_struct, A=A, eigvec=Xr, mode_order=ordq[order], converged=int(conv.sum()),
                       operator_launches=op.n_spmv, phase1_cycles=phase1_cycles, in_band=in_band[order])


def _fp64_block_refine(A, V, k, pol, rig, so, S, tol_rel, checked, verify, maxiter=4, verbose=False, timer=None,
                       band=None):
    """Float64 refinement of the Ritz block V[:k] of a float32 phase: filtered subspace steps
    X <- p(X) with the matrix-free float64 operator (the rigid-body rows projected out per
    structure after each filter), then a float64 Rayleigh-Ritz per structure (Cholesky of the
    segmented Gram matrix, eigh of the projected pencil) and the residuals, until every pair
    meets ``tol_rel`` (every pair inside the band (p_r, r), when given). ``checked``: (theta,
    rel) of the block as it is; ``verify(X)`` gives them for a block X. Returns (theta [S, k],
    rel [S, k]); V[:k] holds the refined vectors."""
    dev = V.device
    f64 = dict(dtype=torch.float64, device=dev)
    op = Operator(A, "mf64", pol=pol, timer=timer)
    n = V.shape[1]
    X = V[:k]
    Y = torch.empty((k, n), **f64)
    h = torch.empty(S * k, **f64)
    theta_h, rel = checked
    for it in range(maxiter):
        worst = np.where(theta_h <= band[1], rel, 0.0).max() if band is not None else rel.max()
        if worst <= tol_rel:
            break
        for q in range(k):
            op.apply(X[q], Y[q])
            for _ in range(2):
                so.rigid(rig, Y[q], Y[q], 1.0)
        Z = A.matfree(Y.T.contiguous()).T.contiguous()
        G = np.empty((S, k, k))
        H = np.empty((S, k, k))
        for q in range(k):
            so.multidot(Y, k, Y[q], h)
            G[:, :, q] = h.view(S, k).cpu().numpy()
            so.multidot(Y, k, Z[q], h)
            H[:, :, q] = h.view(S, k).cpu().numpy()
        G = 0.5 * (G + G.transpose(0, 2, 1))
        H = 0.5 * (H + H.transpose(0, 2, 1))
        L = np.linalg.cholesky(G)
        Li = np.linalg.inv(L)
        C = Li @ H @ Li.transpose(0, 2, 1)
        th, U = np.linalg.eigh(0.5 * (C + C.transpose(0, 2, 1)))
        W = np.ascontiguousarray(Li.transpose(0, 2, 1) @ U)      # [S, k, k]: X_new = W^T Y per structure
        Wd = torch.as_tensor(W, device=dev)
        so.combine(Y, k, Wd, X)
        theta_h, rel = verify(X)
        if verbose:
            print("float64 block refinement %d: worst rel_res %.2e" % (it, rel.max()))
    return theta_h, rel


def solve_assembly(xyz, k_total, rc=8.0, plus_i=1.0, precision="mf64", tol_rel=1e-8, per_component=None,
                   filtered=True, max_degree=5120, maxiter=300, seed=0, verbose=False, factor=3.5, retries=2,
                   timer=None, v0=None, band_factor=1.3, kpm_degree=64, kpm_probes=1, ratio=0.7, filter_right=None,
                   filter_stretch=2.6, deflate=None):
    """Lowest ``k_total`` non-trivial modes of a structure made of disconnected components
    (e.g. subunits of an assembly farther apart than the cutoff). H is block-diagonal, so its
    spectrum is the union of the components' spectra: every component is solved for
    ``per_component`` modes (default ceil(1.15 k_total / C) + 4) in one lockstep run (each with
    its own six rigid-body modes deflated) and the results are merged.

    With the Chebyshev filter (default) the acceptance band [PlusI, r] is chosen to hold
    about 6 C + ``band_factor`` k_total eigenvalues, counted by a Jackson-damped KPM estimate
    (``kpm_degree`` moments of ``kpm_probes`` random vectors; ``filter_right`` gives r
    directly). Only the modes inside it have to converge. The filter is designed for the
    wider interval [PlusI, PlusI + ``filter_stretch`` (r - PlusI)], so that the accepted modes
    sit well inside its pass band (filter values far above those of the rest of the
    spectrum) and converge in few Lanczos steps. The merge takes the lowest
    k_total of them and is checked: fewer than k_total inside the band widens it (x 1.4), and
    a component with all of its ``per_component`` modes inside the band could hide more, so
    the solve is repeated with twice the modes per component; after ``retries`` repetitions
    it raises. Without the filter, every mode must converge and a component whose highest
    computed mode lies below the k_total-th merged eigenvalue triggers the repetition.

    ``precision``: 'mf64' (matrix-free float64), 'f64' (stored float64 blocks: the plain
    float64 solver), 'f32' (stored float32 blocks, no float64 phase) or 'mixed' (float32
    blocks for the bound, the count and the Lanczos phase, then float64 refinement of the
    Ritz block). ``v0``: optional start vector [3N] in the RCM order of the build, or a
    callable (A, component offsets) -> start vector (e.g. a learned warm start).

    Defaults measured on the 2.4M-atom, 1000-mode assembly (60 components): a Lanczos basis
    of 3.5 x the modes per component, the filter stretched 2.6 x beyond the acceptance band
    (degree 12) and the rigid-body rows projected out after each product ('project', see
    solve_segments) took 84 lockstep steps and 1.41 s, against 72 steps of a degree-18 filter
    on a band sized from the reference's undamped count (1.62 s)."""
    from . import chebyshev as CH
    ops.require_gpu()
    dev = torch.device("cuda", torch.cuda.current_device())
    torch.cuda.synchronize()
    t0 = time.time()
    prec = {"mf64": (), "f32": ("f32",), "f64": ("f64",), "mixed": ("f32",)}[precision]
    bulk = "f32" if precision == "mixed" else precision       # operator of the bound and the KPM probes
    A = build_hessian(xyz, rc, plus_i=plus_i, reorder=True, precision=prec, device=dev)
    seg_atoms = components(A)
    C = len(seg_atoms) - 1
    t_build = time.time() - t0
    ks = int(per_component) if per_component else int(np.ceil(1.15 * k_total / C)) + 4
    phases = {"build": t_build}
    t = time.time()
    lo, hi = CH.spectrum_bound(A, precision=bulk, stop_rel=CH.BOUND_STOP_REL) if filtered else (None, None)
    phases["bound"] = time.time() - t
    t = time.time()
    start = v0(A, seg_atoms) if callable(v0) else v0
    if start is not None:
        phases["start_vector"] = time.time() - t
    counter, filter_info = None, None
    for attempt in range(retries + 1):
        pol, band = None, None
        if filtered:
            t = time.time()
            bound = (A.plus_i, hi * (1 + 1e-3) + 1e-12)
            target = 6 * C + band_factor * k_total
            if filter_right is not None:
                right, cnt = float(filter_right), None
            else:
                if counter is None:
                    counter = CH.KPMCounter(A, bound, degree=kpm_degree, probes=kpm_probes, seed=seed,
                                            precision=bulk)
                right = counter.right_for(target, A.plus_i, bound[1])
                cnt = counter.count(right)
            f_right = A.plus_i + float(filter_stretch) * (right - A.plus_i)
            pol = CH.design_filter(A.plus_i, f_right, bound, max_degree, 5, ratio)
            coef = np.asarray(pol.AdjustedCoefficients[:pol.AdjustedDegree + 1], dtype=np.float64)
            band = (float(np.polynomial.chebyshev.chebval((right - pol.cc) / pol.dd, coef)), right)
            phases["filter"] = phases.get("filter", 0.0) + time.time() - t
            filter_info = {"band": [float(A.plus_i), right], "filter_interval": [float(A.plus_i), f_right],
                           "p_at_band_end": band[0], "degree": int(pol.AdjustedDegree),
                           "kpm_count": None if cnt is None else round(float(cnt), 1), "target_count": target,
                           "kpm_products": counter.n_spmv if counter is not None else 0}
        t = time.time()
        r = solve_segments(A, seg_atoms, ks, precision=precision, tol_rel=tol_rel, maxiter=maxiter, seed=seed,
                           factor=factor, verbose=verbose, t0=t0, pol=pol, timer=timer, v0=start, band=band,
                           deflate=deflate)
        phases["solve"] = phases.get("solve", 0.0) + time.time() - t
        if filtered:
            sel = np.argwhere(r.in_band)                           # (component, mode) pairs inside the band
            vals = r.eigval[r.in_band]
            order = np.argsort(vals, kind="stable")[:k_total]
            full = [s for s in range(C) if r.in_band[s].all()]
            if len(vals) >= k_total and not full:
                break
            if attempt == retries:
                raise RuntimeError("the lockstep solve found %d modes inside the band (components %s fill all %d "
                                   "slots); raise per_component or band_factor" % (len(vals), full[:10], ks))
            if len(vals) < k_total:
                band_factor *= 1.4
            if full:
                ks = 2 * ks
            continue
        flat = r.eigval.reshape(-1)
        order = np.argsort(flat, kind="stable")[:k_total]
        sel = np.stack([order // ks, order % ks], 1)
        vals = flat
        cutoff = flat[order[-1]]
        short = [s for s in range(C) if r.eigval[s, -1] <= cutoff]
        if not short:
            break
        if attempt == retries:
            raise RuntimeError("components %s may have more than %d modes below the %d-th eigenvalue; "
                               "raise per_component" % (short[:10], ks, k_total))
        ks = 2 * ks
    pick = sel[order] if filtered else sel
    comp_of, mode_of = pick[:, 0], pick[:, 1]
    torch.cuda.synchronize()
    return BatchResult(eigval=vals[order], component=comp_of, mode_in_component=mode_of,
                       rel_res=r.rel_res[comp_of, mode_of], seconds=time.time() - t0, build_seconds=t_build,
                       components=C, per_component=ks, restarts=r.restarts, n_spmv=r.n_spmv,
                       degree=int(pol.AdjustedDegree) if pol is not None else 0, A=A, seg_atoms=seg_atoms,
                       eigvec=r.eigvec, mode_order=r.mode_order, lam_max=r.lam_max,
                       operator_launches=r.operator_launches, phase1_cycles=r.phase1_cycles, precision=precision,
                       phases=phases, filter_info=filter_info)
