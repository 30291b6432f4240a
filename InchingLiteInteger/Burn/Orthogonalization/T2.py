"""Drop-in for ``InchingLiteInteger/Burn/Orthogonalization/T2.py``: Gram-Schmidt helpers.

* ``T2_vnext_V_MGSvnext(u, V, k=None)`` (reference :64-94): one classical Gram-Schmidt pass
  u -= V (V^T u) over the first k columns of V [n, k] (two device kernels).
* ``OOC2_qnext_Q_ICGSqnext(u, Q)`` (:98-117): iterated CGS (at most 3 passes, stop when the
  norm drops by less than half), then normalise.
* ``OOC2_qnext_Q_MGSqnext(u, Q)`` (:131-137): modified Gram-Schmidt column by column.
Columns of V / Q are vectors (the reference's layout); u is returned.
"""
import torch

from inching_amd import ops


def _rows(V, k):
    """Row-major [k, n] view (or copy) of the first k columns of V [n, >=k]."""
    R = V[:, :k].T
    return R if R.is_contiguous() else R.contiguous()


def _cgs(u, R):
    k = R.shape[0]
    ws = ops.Workspace(k, u.device)
    h = torch.empty(k, dtype=torch.float64, device=u.device)
    nrm = torch.empty(1, dtype=torch.float64, device=u.device)
    ops.multidot(R, k, u, h, ws)
    ops.gemv_update(R, k, h, u, 1.0, nrm2_out=nrm, ws=ws)
    return float(nrm.item())


def T2_vnext_V_MGSvnext(u, V, k=None):
    k = V.shape[1] if k is None else int(k)
    uv = u.reshape(-1)
    if k > 0:
        _cgs(uv, _rows(V, k))
    return u


def OOC2_qnext_Q_ICGSqnext(u, Q):
    uv = u.reshape(-1).clone()
    R = _rows(Q, Q.shape[1])
    r_pre = float(torch.linalg.vector_norm(uv))
    r1 = r_pre
    for _ in range(3):
        r1 = _cgs(uv, R)
        if r1 > r_pre / 2:
            break
        r_pre = r1
    if r1 <= r_pre / 2:
        print("WARNING. still a loss of orthogonality? Something wrong nan?")
    ops.axpby(0.0, uv, 1.0 / r1, uv)
    return uv.reshape(u.shape)


def OOC2_qnext_Q_MGSqnext(u, Q):
    uv = u.reshape(-1).clone()
    R = _rows(Q, Q.shape[1])
    for i in range(R.shape[0]):
        _cgs(uv, R[i:i + 1])
    return uv.reshape(u.shape)
