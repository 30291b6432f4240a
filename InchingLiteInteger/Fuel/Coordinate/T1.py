"""Drop-in for ``InchingLiteInteger/Fuel/Coordinate/T1.py`` of the reference: 3DRCM ordering
and per-row-batch column islands.

``X_KdCuthillMckeeOrder`` (reference :171-323) and ``X_KdUngappedMinMaxNeighbor``
(reference :331-481) keep their names, arguments and return values. Coordinates are in nm
and ``rc_Gamma`` in Angstrom as in the reference (the cutoff is divided by 10). Neighbour
sets come from the device cell list (``inch_neighbors_celllist``); the permutation from the
level-synchronous device RCM (``inch_rcm_device``), which reproduces the reference's BFS,
child ordering and reversal exactly, with ties in the seed order broken by a stable argsort
(pass ``User_SeedOrder`` to reproduce any other tie order, e.g. ``numpy.argsort(degree)``).

Periodic images in x and y (``User_DictCharmmGuiPbc``, the CHARMM-GUI box dictionary with
``RectBox_Xsize``, ``RectBox_Ysize`` and the ``X`` / ``Y`` ranges, in nm like X) follow
``X_cKDTreePbcXy`` (reference :67-161): j is a neighbour of i when any of the nine x/y
images of j is within the cutoff. The reference visits those neighbours in the order its
k-d tree returns the concatenated image lists (an internal order of scipy); here each row
is the sorted set of such j, so the periodic permutation follows the same BFS rules on that
set and is not bit-identical to the reference's.
"""
import numpy as np
import torch

from inching_amd import ops
from inching_amd.hessian import _to_device_xyz


def pbc_box(User_DictCharmmGuiPbc, rc_nm):
    """(box, origin) of a CHARMM-GUI box dictionary, with the reference's size assertions."""
    d = User_DictCharmmGuiPbc
    assert (d['RectBox_Xsize'] > rc_nm), "ABORTED. The PBC box size X is smaller than rc gamma."
    assert (d['RectBox_Ysize'] > rc_nm), "ABORTED. The PBC box size Y is smaller than rc gamma."
    return (float(d['RectBox_Xsize']), float(d['RectBox_Ysize'])), (float(d["X"][0]), float(d["Y"][0]))


def _neighbours_device(X, rc_nm, User_DictCharmmGuiPbc=None):
    ops.require_gpu()
    dev = torch.device("cuda", torch.cuda.current_device())
    xyz = _to_device_xyz(X, dev)
    if User_DictCharmmGuiPbc is None:
        return ops.neighbours(xyz, rc_nm)
    box, origin = pbc_box(User_DictCharmmGuiPbc, rc_nm)
    ip, ix, _ = ops.neighbours_pbc(xyz, rc_nm, box, origin)
    return ip, ix


def _neighbours(X, rc_nm, User_DictCharmmGuiPbc=None):
    ip, ix = _neighbours_device(X, rc_nm, User_DictCharmmGuiPbc)
    return ip.cpu().numpy(), ix.cpu().numpy()


def X_KdCuthillMckeeOrder(X, rc_Gamma=15.0, Reverse=True, ReturnStat=False, User_maxleafsize=1000,
                          User_DictCharmmGuiPbc=None, User_SeedOrder=None):
    """(order, undo) of the 3DRCM permutation; with ReturnStat also mean and std of the
    neighbour count (self included)."""
    ip, ix = _neighbours_device(X, rc_Gamma / 10.0, User_DictCharmmGuiPbc)
    order, undo = ops.rcm_device(ip, ix, seed_order=User_SeedOrder, reverse=Reverse)
    order, undo = order.cpu().numpy(), undo.cpu().numpy()
    if ReturnStat:
        deg = np.diff(ip.cpu().numpy())
        return order, undo, np.mean(deg), np.std(deg)
    return order, undo


def ungapped_islands(indptr, indices, n_atoms, maxleafsize=1000, gap_size=100):
    """Per batch the sorted unique neighbour columns split at gaps wider than ``gap_size``,
    with the reference's exceptions (the first five and the fifth-last difference are never
    gaps; a gap whose left island reaches the batch's diagonal block is not cut)."""
    head = ops.batch_head(n_atoms, maxleafsize)
    out = {}
    total = 0
    for b in range(len(head) - 1):
        start, end = int(head[b]), int(head[b + 1])
        cols = indices[indptr[start]:indptr[end]]
        total += cols.size
        uniq = np.unique(cols)
        nu = uniq.size
        cand = np.nonzero(np.diff(uniq) > gap_size)[0]
        gse = [int(uniq[0])]
        height = end - start
        for iii in cand:
            if iii == nu - 5 or iii < 5:
                continue
            if (gse[-1] + 5 >= start) and (int(uniq[iii]) + 1 <= start + height + 5):
                continue
            gse.extend([int(uniq[iii]) + 1, int(uniq[iii + 1])])
        gse.append(int(uniq[-1]) + 1)
        out[b] = [(gse[t], gse[t + 1]) for t in range(0, len(gse), 2)]
    half_nnz = (((total - n_atoms) / 2) + n_atoms) * 9
    return out, half_nnz


def X_KdUngappedMinMaxNeighbor(X, rc_Gamma=15.0, maxleafsize=1000, CollectStat=False, User_ReturnHalfNnz=False,
                               User_GapSize=100, User_DictCharmmGuiPbc=None, SliceForm=True):
    """{batch: [(start, end), ...]} column islands [, HalfNnz].

    The islands are in slice form (end exclusive) whatever ``SliceForm`` says: the reference
    returns its slice-form dictionary in both cases (``:470-481``); ``SliceForm`` only shaped a
    per-batch (min, max) dictionary that it never returned."""
    ip, ix = _neighbours(X, rc_Gamma / 10.0, User_DictCharmmGuiPbc)
    n = ip.shape[0] - 1
    islands, half = ungapped_islands(ip, ix, n, maxleafsize=maxleafsize, gap_size=User_GapSize)
    if User_ReturnHalfNnz:
        return islands, half
    return islands
