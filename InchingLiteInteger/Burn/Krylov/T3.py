"""Drop-in for ``InchingLiteInteger/Burn/Krylov/T3.py`` of the reference: matvec closures.

``OOC2_HalfMemS_v_KrylovAv_VOID(A, A_diag)`` (reference :54-165) and
``OOC2_FullMemS_v_KrylovAv_VOID(A, A_diag)`` (:169-239) return ``aux(A, v, u)``, which
writes ``u = A v`` in place (the reference computes L v + L^T v - A_diag v with two
cuSPARSE calls). Here one launch of the fused block SpMV kernel does it; u and v are
torch tensors on the GPU.
"""
from inching_amd import ops

from ..._common import require_block_matrix


def _closure(A, A_diag):
    require_block_matrix(A)

    def aux(A_, v, u):
        M = require_block_matrix(A_)
        prec = "f64" if M.vals64 is not None else "f32"
        vv = v.reshape(-1)
        uu = u.reshape(-1)
        if not vv.is_contiguous() or not uu.is_contiguous():
            raise ValueError("v and u must be contiguous")
        if vv.data_ptr() == uu.data_ptr():
            raise ValueError("u must not alias v")
        ops.bsr_spmv(M.n_atoms, M.rowptr, M.col, M.vals(prec), vv, uu, lanes=M.lanes)
    return aux


def OOC2_HalfMemS_v_KrylovAv_VOID(A, A_diag):
    return _closure(A, A_diag)


def OOC2_FullMemS_v_KrylovAv_VOID(A, A_diag):
    return _closure(A, A_diag)
