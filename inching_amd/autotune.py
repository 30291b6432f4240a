"""Launch-shape autotuner for the operator kernels.

The block SpMV takes ``lanes`` (lanes of a 64-wide wavefront per block row) and the
matrix-free ANM product takes ``lanes*10 + unroll``. The best shape depends on the mean
neighbour count (row length), the atom count (grid size against 256 CUs) and the value
precision, so each candidate is timed with HIP events on the matrix at hand (a few
launches each) and the winner is cached per (kernel, precision, log2 atoms, row length).
The cache lives in the process; ``INCHING_TUNE_CACHE`` names an optional JSON file that
persists it.
"""
import json
import math
import os

import torch

from . import ops

BSR_LANES = (4, 8, 16, 32, 64)
# Times within NOISE of the fastest count as ties, resolved by a fixed preference order
# (candidates listed first win), so that the choice does not flip between runs on timing
# noise: block SpMV shapes with 16-bit columns and XCD-contiguous rows first (the lowest
# measured HBM traffic), wide row groups before narrow ones; cluster kernels before per-row.
NOISE = 0.02
BSR_PREFERENCE = tuple(1200 + g for g in (32, 64, 16, 8, 4)) + tuple(200 + g for g in (32, 64, 16, 8, 4)) + \
    (32, 64, 16, 8, 4)
CLUSTER_PREFERENCE = (6001, 6104, 6002, 6108, 6004, 6116, 4001, 4104, 4002, 4108, 4004, 4116)
MATFREE_CONFIGS = (41, 42, 81, 82, 161, 162, 321, 322)
CLUSTER_CONFIGS = (1, 2, 4, 104, 108, 116)   # list entries per lane iteration; 100 + clusters per wave (persistent)
# From this many atoms on, a matrix with the cluster layout times only the cluster kernels:
# their best shape was faster than the best per-row shape by 1.22x (30k-atom globule), 1.50x
# (300k), 1.39x (512-globule batch, 2.7M atoms) and 1.45x (2.4M-atom assembly), while the
# per-row shapes won below (0.85x at 1,200 atoms, 0.94x at 5,000). Timing the 32 per-row
# shapes of a 2.7M-atom batch took about 0.35 s. Such matrices skip the persistent cluster
# shapes too, which lost to the plain ones there (0.101 vs 0.092 ms at 300k atoms, 0.894 vs
# 0.815 ms on the assembly) and took 47 ms of tuning on the batch.
CLUSTER_ONLY_ATOMS = 20000

_cache = {}
_loaded = False


def _path():
    return os.environ.get("INCHING_TUNE_CACHE")


def _load():
    global _loaded
    if _loaded:
        return
    _loaded = True
    p = _path()
    if p and os.path.exists(p):
        try:
            with open(p) as f:
                _cache.update(json.load(f))
        except (OSError, ValueError):
            pass


def _save():
    p = _path()
    if p:
        os.makedirs(os.path.dirname(os.path.abspath(p)), exist_ok=True)
        with open(p, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)


def key(kind, precision, A):
    deg = A.nnzb / max(1, A.n_atoms)
    return "%s/%s/n2^%d/deg%d" % (kind, precision, int(round(math.log2(max(2, A.n_atoms)))), int(round(deg / 8.0)) * 8)


def _time(fn, reps):
    for _ in range(2):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def best_config(A, precision, reps=5):
    """Lanes (block SpMV) or lanes*10+unroll (matrix-free) for this matrix and precision."""
    _load()
    kind = "matfree" if precision == "mf64" else "bsr"
    if kind == "matfree" and getattr(A, "images", None) is not None:
        return 161            # periodic images: one kernel shape, nothing to tune
    k = key(kind, precision, A)
    if k in _cache:
        return int(_cache[k]["best"])
    n = A.shape[0]
    vdt = torch.float32 if precision == "f32" else torch.float64
    x = torch.randn(n, dtype=vdt, device=A.device)
    z = torch.randn_like(x)
    y = torch.empty_like(x)
    acc = torch.zeros_like(x)
    times = {}
    if kind == "matfree":
        # lanes*10+unroll, + 1000 for 16-bit column offsets, + 2000 for XCD-contiguous rows,
        # + 4000 for the cluster layout
        clusters = A.clusters() is not None
        large = A.n_atoms >= CLUSTER_ONLY_ATOMS
        codes = ()
        if not clusters or not large:
            flags = (0, 2000) + ((1000, 3000) if A.col16() is not None else ())
            codes = tuple(f + c for f in flags for c in MATFREE_CONFIGS)
        if clusters:
            codes += tuple(f + c for f in (4000, 6000) for c in CLUSTER_CONFIGS if not (large and c > 100))
        for cfg in codes:
            times[cfg] = _time(lambda: A.matfree_spmv(x, y, s1=0.1, c0=1.0, z=z, c2=1.0, acc=acc, mu=0.1, config=cfg),
                               reps)
        order = CLUSTER_PREFERENCE + tuple(c for c in codes if c < 4000)
    else:
        order = BSR_PREFERENCE
        # int32 columns: code = lanes; 16-bit column offsets: code = 200 + lanes
        vals = A.vals(precision)
        for lanes in BSR_LANES:
            times[lanes] = _time(lambda: ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, vals, x, y, s1=0.1, c0=1.0, z=z,
                                                      c2=1.0, acc=acc, mu=0.1, lanes=lanes), reps)
        c16 = A.col16()
        if c16 is not None:
            for lanes in BSR_LANES:
                for base, arg in ((200, lanes), (1200, lanes + 1000)):       # 1200+: XCD-contiguous rows
                    times[base + lanes] = _time(lambda: ops.bsr_spmv(A.n_atoms, A.rowptr, A.col, vals, x, y, s1=0.1,
                                                                     c0=1.0, z=z, c2=1.0, acc=acc, mu=0.1, lanes=arg,
                                                                     col16=c16), reps)
    best = _pick(times, order)
    _cache[k] = {"best": int(best), "ms": {str(c): round(t, 5) for c, t in times.items()}}
    _save()
    return int(best)


def _pick(times, order):
    """The first candidate in ``order`` whose time is within NOISE of the fastest."""
    tmin = min(times.values())
    for c in order:
        if c in times and times[c] <= tmin * (1.0 + NOISE):
            return c
    return min(times, key=times.get)


def bsr_launch(A, code):
    """(lanes argument, col16 or None) of a block-SpMV config code: lanes (int32 columns),
    200 + lanes (16-bit columns), 1200 + lanes (16-bit columns, XCD-contiguous rows)."""
    code = int(code)
    if code >= 1200:
        return code - 1200 + 1000, A.col16()
    if code >= 200:
        return code - 200, A.col16()
    return code, None


def table():
    _load()
    return dict(_cache)
