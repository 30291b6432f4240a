"""3DRCM restatement (oracle).

Follows ``InchingLiteInteger/Fuel/Coordinate/T1.py:171-323`` (``X_KdCuthillMckeeOrder``)
and ``Tex/main.tex:63-146`` (Alg. 1):

* degree[i] = |ball(i)| including i itself (:236-245);
* seeds are taken in the order ``inds = argsort(degree)`` (:254); the reference uses
  numpy's default (unstable) kind. This restatement pins the tie rule to a *stable*
  argsort and also accepts an explicit seed order so any other tie rule can be
  reproduced (``seed_order=np.argsort(degree)`` reproduces the reference on a given
  numpy build);
* BFS per component; a dequeued node's neighbours are visited in *descending*
  index order (``query_ball_point(..., return_sorted=True)[::-1]``, :276-277);
* the newly appended children of that node are stably insertion-sorted by
  degree (:294-304);
* the final order is reversed (:320-321); the second output is
  ``argsort(order)`` (the undo permutation).
"""
import numpy as np


def rcm_order(indptr, indices, seed_order=None, reverse=True):
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    n = indptr.shape[0] - 1
    degree = np.diff(indptr)
    if seed_order is None:
        inds = np.argsort(degree, kind="stable")
    else:
        inds = np.asarray(seed_order, dtype=np.int64).copy()
    visited = np.zeros(n, dtype=bool)
    order = np.zeros(n, dtype=np.int64)
    N = 0
    for zz in range(n):
        seed = inds[zz]
        if visited[seed]:
            continue
        order[N] = seed
        N += 1
        visited[seed] = True
        level_start, level_end = N - 1, N
        while level_start < level_end:
            for ii in range(level_start, level_end):
                i = order[ii]
                n_old = N
                row = indices[indptr[i]:indptr[i + 1]]
                for j in row[::-1]:
                    if not visited[j]:
                        visited[j] = True
                        order[N] = j
                        N += 1
                if N - n_old > 1:
                    kids = order[n_old:N]
                    # stable sort by degree == the reference's strict-< insertion sort
                    order[n_old:N] = kids[np.argsort(degree[kids], kind="stable")]
            level_start, level_end = level_end, N
        if N == n:
            break
    if reverse:
        out = order[::-1].copy()
    else:
        out = order
    return out.astype(np.int32), np.argsort(out).astype(np.int64)


def rcm_order_levels(indptr, indices, seed_order=None, reverse=True):
    """Level-synchronous form of :func:`rcm_order` (the order the device kernels build):
    level L+1 = unvisited neighbours of level L, each owned by the first level-L node
    adjacent to it, sorted by (owner position, degree, -index). Equal to rcm_order on every
    graph; kept as a restatement to check that equivalence on the CPU."""
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    n = indptr.shape[0] - 1
    degree = np.diff(indptr)
    inds = np.argsort(degree, kind="stable") if seed_order is None else np.asarray(seed_order, dtype=np.int64)
    pos = np.full(n, -1, dtype=np.int64)
    order = []
    for seed in inds:
        if pos[seed] >= 0:
            continue
        pos[seed] = len(order)
        order.append(seed)
        level = [seed]
        while level:
            owner = {}
            for p_node in level:
                p = pos[p_node]
                for j in indices[indptr[p_node]:indptr[p_node + 1]]:
                    if pos[j] < 0 and (j not in owner or owner[j] > p):
                        owner[j] = p
            nxt = sorted(owner, key=lambda j: (owner[j], degree[j], -j))
            for j in nxt:
                pos[j] = len(order)
                order.append(j)
            level = nxt
    out = np.asarray(order[::-1] if reverse else order, dtype=np.int64)
    return out.astype(np.int32), np.argsort(out).astype(np.int64)


def bandwidth(indptr, indices, order):
    """Max |p(i) - p(j)| over neighbour pairs after permuting by ``order``."""
    indptr = np.asarray(indptr, dtype=np.int64)
    pos = np.empty(len(order), dtype=np.int64)
    pos[np.asarray(order, dtype=np.int64)] = np.arange(len(order))
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    if rows.size == 0:
        return 0
    return int(np.max(np.abs(pos[rows] - pos[np.asarray(indices, dtype=np.int64)])))
