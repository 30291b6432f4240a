"""Drop-in for ``InchingLiteInteger/Burn/Orthogonalization/T3.py``: the Hotelling (EED) term.

``T3_QHotelling_x_Ax_HotelledAx(QHotelling, x, res, HotellingShift=10.0)`` (reference
:62-164) adds ``HotellingShift * Q^T (Q x)`` to ``res`` in place and returns it; Q may be
dense or sparse (m x n), x and res are float64 GPU tensors.
"""
import torch

from inching_amd import ops

from ..._common import dense_Q


def T3_QHotelling_x_Ax_HotelledAx(QHotelling, x, res, HotellingShift=10.0):
    xv = x.reshape(-1)
    rv = res.reshape(-1)
    Q = dense_Q(QHotelling, xv.device)
    m = Q.shape[0]
    h = torch.empty(m, dtype=torch.float64, device=xv.device)
    ws = ops.Workspace(m, xv.device)
    ops.multidot(Q, m, xv.contiguous(), h, ws)
    ops.gemv_update(Q, m, h, rv, scale=-float(HotellingShift))
    return res
