"""Drop-in for ``InchingLiteInteger/Burn/Orthogonalization/T1.py``: row Gram-Schmidt.

``R_OrthogonalizeAllBefore(V, ix, beforeindex)`` (reference :56-64): V[ix] is projected
against the rows V[:beforeindex] and normalised, in place; V (rows are vectors) is returned.
"""
import torch

from inching_amd import ops


def R_OrthogonalizeAllBefore(V, ix, beforeindex):
    ws = ops.Workspace(max(1, beforeindex), V.device)
    nrm = torch.empty(1, dtype=torch.float64, device=V.device)
    if beforeindex > 0:
        h = torch.empty(beforeindex, dtype=torch.float64, device=V.device)
        ops.multidot(V, beforeindex, V[ix], h, ws)
        ops.gemv_update(V, beforeindex, h, V[ix], 1.0, nrm2_out=nrm, ws=ws)
    else:
        ops.gemv_update(V[ix:ix + 1], 0, None, V[ix], 0.0, nrm2_out=nrm, ws=ws)
    ops.scale_div(V[ix], nrm, V[ix])
    return V
