"""Device operator for the eigensolvers: u = op(v) with op one of

* A v                                   (``Burn/Krylov/T3.py:54-239``)
* p(A) v, p the Chebyshev filter        (``Burn/PolynomialFilters/T2.py:28-92``)
* either of the above + shift Q^T Q v   (Hotelling / EED, ``Burn/Orthogonalization/T3.py:62-164``)

Precision policy: the block values and the Chebyshev recurrence vectors are float32 with
float32 accumulation in ``precision='f32'`` (mixed) mode and float64 in ``'f64'`` mode;
``'mf64'`` uses no stored values at all: the float64 ANM product is recomputed from the
float64 coordinates and the neighbour pattern (4 B of index traffic per neighbour instead
of 36-72 B of block values). v and u (the Lanczos basis side) are always float64. One
filtered matvec of degree m is m launches of the fused SpMV + three-term-recurrence kernel.

``timer``: optional list; when set, every operator launch is bracketed by HIP events on
the current stream and the event pairs are appended (kernel-time accounting).

With the matrix-free float64 operator on the cluster layout, a filtered application runs
its recurrence on records (``ops.anm_records``): the input is packed once into cluster order
next to the coordinates, the degree products read and write records, and the accumulated
result is unpacked once (``records=False`` or ``INCHING_RECORDS=0`` keeps the vectors in the
matrix's order).
"""
import os

import torch

from . import ops

RECORDS = os.environ.get("INCHING_RECORDS", "1") != "0"


class Operator:
    def __init__(self, A, precision="f64", pol=None, Q=None, shift=0.0, lanes=None, timer=None, tune=True,
                 records=None):
        if precision not in ("f32", "f64", "mf64"):
            raise ValueError("precision must be 'f32', 'f64' or 'mf64'")
        self.A = A
        self.precision = precision
        self.timer = timer
        if precision == "mf64":
            self.vals = None
        else:
            A.ensure(precision)
            self.vals = A.vals(precision)
        self.n = A.shape[0]
        self.pol = pol
        self.Q = Q
        self.shift = float(shift)
        use_rec = RECORDS if records is None else bool(records)
        rec_ok = (pol is not None and precision == "mf64" and use_rec and getattr(A, "images", None) is None
                  and (lanes is None or int(lanes) // 1000 & 4) and A.records() is not None)
        if lanes is not None:
            self.lanes = int(lanes)
        elif rec_ok:
            self.lanes = 6001 if A.clusters() is not None else 161   # plain products only; no tuning needed
        elif tune:
            from .autotune import best_config
            self.lanes = best_config(A, precision)
        else:
            self.lanes = 161 if precision == "mf64" else A.lanes
        self.blanes, self.c16 = self.lanes, None
        if precision != "mf64":
            from .autotune import bsr_launch
            self.blanes, self.c16 = bsr_launch(A, self.lanes)
        self.n_spmv = 0
        self.n_apply = 0
        dev = A.device
        vdt = torch.float32 if precision == "f32" else torch.float64
        self.vdt = vdt
        if pol is not None:
            self._vk = torch.empty(self.n, dtype=vdt, device=dev)
            self._vkm1 = torch.empty(self.n, dtype=vdt, device=dev)
            self._vkp1 = torch.empty(self.n, dtype=vdt, device=dev)
            self._acc = torch.empty(self.n, dtype=vdt, device=dev)
            self.mu = [float(x) for x in pol.AdjustedCoefficients[:pol.AdjustedDegree + 1]]
            self.degree = int(pol.AdjustedDegree)
        self.rec = None
        if rec_ok:
            self.rec = A.records()
            self._acc_rec = torch.empty(3 * A.clusters()[0].numel(), dtype=torch.float64, device=dev)
        if Q is not None:
            if Q.dtype != torch.float64 or Q.shape[1] != self.n:
                raise ValueError("Q must be float64 [m, 3N]")
            self.Q = Q.contiguous()
            self._h = torch.empty(Q.shape[0], dtype=torch.float64, device=dev)
            self._ws = ops.Workspace(Q.shape[0], dev)

    def spmv(self, x, y, _timed=True, **epi):
        timed = self.timer is not None and _timed
        if timed:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
        if self.vals is None:
            self.A.matfree_spmv(x, y, config=self.lanes, **epi)
        else:
            ops.bsr_spmv(self.A.n_atoms, self.A.rowptr, self.A.col, self.vals, x, y, lanes=self.blanes,
                         col16=self.c16, **epi)
        if timed:
            e1.record()
            epi = "cheb" if (epi.get("acc") is not None or epi.get("z") is not None) else "plain"
            self.timer.append((self.precision, epi, e0, e1, 1))
        self.n_spmv += 1

    def _filtered_records(self, v64, u64):
        """u = sum_k mu_k T_k((A - cc)/dd) v with the recurrence on records: v packed into
        cluster order, three record buffers in turn (T_k, T_{k-1} read, T_{k+1} written),
        the accumulator in slot order unpacked into u at the end."""
        pol = self.pol
        dd, cc = float(pol.dd), float(pol.cc)
        recs, rent = self.rec
        atoms, ptr = self.A.clusters()[:2]
        acc = self._acc_rec
        plus_i = self.A.plus_i
        cur, prev = recs[0], None
        free = [recs[1], recs[2]]
        ops.rec_pack(atoms, v64, cur, acc, self.mu[0])
        if self.timer is not None:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
        for k in range(1, self.degree + 1):
            out = free.pop(0)
            if k == 1:
                ops.matfree_rec(cur, None, out, ptr, rent, plus_i, s1=1.0 / dd, c0=cc, acc=acc, mu=self.mu[1])
            else:
                ops.matfree_rec(cur, prev, out, ptr, rent, plus_i, s1=2.0 / dd, c0=cc, c2=1.0, acc=acc,
                                mu=self.mu[k])
            if prev is not None:
                free.append(prev)
            prev, cur = cur, out
        if self.timer is not None:
            e1.record()
            self.timer.append(("mf64rec", "cheb", e0, e1, self.degree))
        ops.rec_unpack(atoms, acc, u64)
        self.n_spmv += self.degree

    def filtered(self, v64, u64):
        """u = sum_k mu_k T_k((A - cc)/dd) v.

        With a float64 operator the recurrence starts from v itself (it is only read) and,
        when u does not share v's memory, accumulates straight into u; otherwise v is
        converted into a scratch vector and the accumulator is converted into u at the end."""
        if (self.rec is not None and v64.dtype == torch.float64 and u64.dtype == torch.float64
                and v64.is_contiguous() and u64.is_contiguous()):
            self._filtered_records(v64, u64)
            return
        pol = self.pol
        dd, cc = float(pol.dd), float(pol.cc)
        scratch = [self._vk, self._vkm1, self._vkp1]
        f64 = self.vdt == torch.float64
        if f64 and v64.dtype == torch.float64 and v64.is_contiguous():
            cur = v64
        else:
            cur = scratch[0]
            ops.convert(v64, cur)
        direct = (f64 and u64.dtype == torch.float64 and u64.is_contiguous()
                  and u64.data_ptr() != v64.data_ptr())
        acc = u64 if direct else self._acc
        free = [b for b in scratch if b is not cur]
        ops.axpby(self.mu[0], v64, 0.0, acc)
        # the degree operator launches run back to back: one event pair times all of them
        if self.timer is not None:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
        prev = None
        for k in range(1, self.degree + 1):
            out = free.pop(0)
            if k == 1:
                self.spmv(cur, out, _timed=False, s1=1.0 / dd, c0=cc, acc=acc, mu=self.mu[1])
            else:
                self.spmv(cur, out, _timed=False, s1=2.0 / dd, c0=cc, z=prev, c2=1.0, acc=acc, mu=self.mu[k])
            if prev is not None and prev is not v64:
                free.append(prev)            # T_{k-2} v is no longer needed (v itself is never written)
            prev, cur = cur, out
        if self.timer is not None:
            e1.record()
            self.timer.append((self.precision, "cheb", e0, e1, self.degree))
        if not direct:
            ops.convert(acc, u64)

    def apply(self, v64, u64):
        self.n_apply += 1
        if self.pol is None:
            self.spmv(v64, u64)
        else:
            self.filtered(v64, u64)
        if self.Q is not None and self.shift != 0.0:
            m = self.Q.shape[0]
            ops.multidot(self.Q, m, v64, self._h, self._ws)
            ops.gemv_update(self.Q, m, self._h, u64, scale=-self.shift)
        return u64

    def plain(self, v64, u64):
        """A v without filter and without the Hotelling term."""
        self.spmv(v64, u64)
        return u64
