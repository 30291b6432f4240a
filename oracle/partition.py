"""Row-batch partition and ungapped column islands (oracle).

``partition_leaves`` / ``batch_head`` restate ``InchingLiteInteger/util.py:563-587``
(``GetPartitionTree`` + ``FlattenPartitionTree``) and the batch_head construction of
``InchingLiteInteger/Fuel/T1.py:145-153``: recursive halving of n into
``range(0, k+1)`` and ``range(k+1, n)`` with k = floor(n/2) until a piece has at most
``maxleafsize`` elements; the leaf sizes are sorted descending and cumulated.

``ungapped_islands`` restates ``X_KdUngappedMinMaxNeighbor``
(``InchingLiteInteger/Fuel/Coordinate/T1.py:331-481``) on a precomputed neighbour CSR.
"""
import numpy as np


def partition_leaves(n, maxleafsize):
    out = []
    stack = [n]
    # depth-first, left piece first (the order of FlattenPartitionTree)
    while stack:
        m = stack.pop()
        if m <= maxleafsize:
            out.append(m)
            continue
        k = m // 2
        stack.append(m - (k + 1))
        stack.append(k + 1)
    return out


def batch_head(n, maxleafsize):
    leaves = sorted(partition_leaves(n, maxleafsize))[::-1]
    head = [0]
    for s in leaves:
        head.append(head[-1] + s)
    return np.asarray(head, dtype=np.int64)


def ungapped_islands(indptr, indices, maxleafsize=100, gap_size=100):
    """Return ({batch: [(c0, c1), ...]}, HalfNnz) exactly as the reference builds them."""
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    n = indptr.shape[0] - 1
    head = batch_head(n, maxleafsize)
    out = {}
    total = 0
    for b in range(len(head) - 1):
        start, end = int(head[b]), int(head[b + 1])
        cols = indices[indptr[start]:indptr[end]]
        total += cols.shape[0]
        batch_height = end - start
        uniq = np.unique(cols)
        cmin, cmax = int(uniq[0]), int(uniq[-1])
        diffs = np.diff(uniq)
        gse = [cmin]
        nuniq = uniq.shape[0]
        for iii in range(diffs.shape[0]):
            xxx = int(diffs[iii])
            if iii == nuniq - 5:
                continue
            if iii < 5:
                continue
            if xxx > gap_size:
                if (gse[-1] + 5 >= start) and (int(uniq[iii]) + 1 <= (start + batch_height + 5)):
                    continue
                gse.extend([int(uniq[iii]) + 1, int(uniq[iii + 1])])
        gse.append(cmax + 1)
        out[b] = [(gse[t], gse[t + 1]) for t in range(0, len(gse), 2)]
    half_nnz = (((total - n) / 2) + n) * 9
    return out, half_nnz
