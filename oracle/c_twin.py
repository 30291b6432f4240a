"""ctypes wrapper of the plain-C oracle (``oracle/c/rcm_oracle.c``): neighbour CSR and 3DRCM.

Test infrastructure, like the rest of ``oracle``. ``build()`` compiles the C file with gcc
(``-O2 -ffp-contract=off`` so the float64 distance rule is evaluated without contraction)
into ``oracle/_build/liboracle_c.so``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "rcm_oracle.c")
OUT = os.path.join(HERE, "_build", "liboracle_c.so")
_LIB = None


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c99", SRC, "-o", OUT, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return OUT


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(OUT) or os.path.getmtime(OUT) < os.path.getmtime(SRC):
            build()
        L = C.CDLL(OUT)
        L.oracle_neighbours.restype = C.c_int64
        L.oracle_neighbours.argtypes = [C.c_int64, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
        L.oracle_rcm.restype = C.c_int
        L.oracle_rcm.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        _LIB = L
    return _LIB


def neighbour_csr(X, rc):
    X = np.ascontiguousarray(X, dtype=np.float64)
    n = X.shape[0]
    indptr = np.zeros(n + 1, dtype=np.int64)
    L = lib()
    tot = L.oracle_neighbours(n, X.ctypes.data, float(rc), indptr.ctypes.data, None)
    if tot < 0:
        raise MemoryError("oracle_neighbours")
    indices = np.zeros(max(int(tot), 1), dtype=np.int32)
    L.oracle_neighbours(n, X.ctypes.data, float(rc), indptr.ctypes.data, indices.ctypes.data)
    return indptr, indices[:int(tot)]


def rcm_order(indptr, indices, seed_order=None, reverse=True):
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    n = indptr.size - 1
    order = np.zeros(n, dtype=np.int32)
    undo = np.zeros(n, dtype=np.int64)
    so = None if seed_order is None else np.ascontiguousarray(seed_order, dtype=np.int64)
    st = lib().oracle_rcm(n, indptr.ctypes.data, indices.ctypes.data, None if so is None else so.ctypes.data,
                          1 if reverse else 0, order.ctypes.data, undo.ctypes.data)
    if st != 0:
        raise MemoryError("oracle_rcm")
    return order, undo
