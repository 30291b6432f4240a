"""Typed wrappers of the C ABI on torch tensors (device memory + the current HIP stream).

Every wrapper checks shapes and dtypes on the host before launching, so a kernel never
sees a buffer smaller than its grid assumes.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import F32, F64, InchingError, call  # noqa: F401  (InchingError re-exported)

_DT = {torch.float32: F32, torch.float64: F64}


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("inching_amd: a ROCm GPU is required for device kernels (torch.cuda.is_available() is False)")


def stream_ptr(stream=None):
    """hipStream_t of ``stream``, or of torch's current stream (the raw accessor avoids
    constructing a Stream object, which costs several microseconds per launch)."""
    if stream is not None:
        return stream.cuda_stream
    return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())


def P(t):
    """Raw device/host pointer of a tensor as an int (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def dt(t):
    try:
        return _DT[t.dtype]
    except KeyError:
        raise TypeError("unsupported dtype %s" % t.dtype)


def _check_dev(*ts):
    for t in ts:
        if t is not None:
            if not t.is_cuda:
                raise ValueError("tensor must live on the GPU")
            if not t.is_contiguous():
                raise ValueError("tensor must be contiguous")


# ------------------------------------------------------------------ host ----

def synth_globule(n, radius, dmin, seed, max_attempts=0):
    out = np.empty((int(n), 3), dtype=np.float64)
    call("inch_synth_globule", int(n), float(radius), float(dmin), int(seed) & 0xFFFFFFFFFFFFFFFF,
         int(max_attempts), out.ctypes.data_as(C.c_void_p))
    return out


def rcm_host(indptr, indices, seed_order=None, reverse=True):
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    n = indptr.shape[0] - 1
    order = np.empty(n, dtype=np.int32)
    undo = np.empty(n, dtype=np.int64)
    so = None
    if seed_order is not None:
        so = np.ascontiguousarray(seed_order, dtype=np.int64)
        if so.shape != (n,):
            raise ValueError("seed_order must have shape (n,)")
    call("inch_rcm_host", n, indptr.ctypes.data_as(C.c_void_p), indices.ctypes.data_as(C.c_void_p),
         so.ctypes.data_as(C.c_void_p) if so is not None else None, 1 if reverse else 0,
         order.ctypes.data_as(C.c_void_p), undo.ctypes.data_as(C.c_void_p))
    return order, undo


def rcm_device(indptr, indices, seed_order=None, reverse=True, stream=None, components=False):
    """3DRCM on the device: (order int32 [n], undo int64 [n]) device tensors, bit-exact to
    rcm_host. indptr int64 / indices int32 device tensors (unique columns per row).
    seed_order: optional seed order (default: stable argsort of the degree, ties in index
    order as numpy's stable argsort, sorted on the device). ``components``: also return the
    start offsets of the connected components in the output order (numpy int64 [C + 1])."""
    n = indptr.numel() - 1
    dev = indptr.device
    if n <= 0:
        raise ValueError("empty graph")
    if seed_order is None:
        seeds = torch.argsort(indptr[1:] - indptr[:-1], stable=True)
    else:
        seeds = torch.as_tensor(np.ascontiguousarray(seed_order, dtype=np.int64), device=dev).reshape(-1)
        if seeds.numel() != n:
            raise ValueError("seed_order must have shape (n,)")
    order = torch.empty(n, dtype=torch.int32, device=dev)
    undo = torch.empty(n, dtype=torch.int64, device=dev)
    wb = int(call("inch_rcm_device_work_size", n))
    work = torch.empty(wb, dtype=torch.uint8, device=dev)
    cap = n + 1
    comp = np.zeros(cap, dtype=np.int64) if components else None
    ncomp = C.c_int64(0)
    call("inch_rcm_device", n, P(indptr), P(indices), P(seeds), 1 if reverse else 0, P(order), P(undo), P(work), wb,
         comp.ctypes.data_as(C.c_void_p) if components else None, cap, C.byref(ncomp), stream_ptr(stream))
    if components:
        return order, undo, comp[:ncomp.value + 1].copy()
    return order, undo


def batch_head(n, maxleaf):
    cap = 2 * (int(n) // max(1, int(maxleaf) // 2) + 2) + 8
    head = np.empty(cap, dtype=np.int64)
    nb = C.c_int64(0)
    call("inch_batch_head", int(n), int(maxleaf), head.ctypes.data_as(C.c_void_p), cap, C.byref(nb))
    return head[:nb.value + 1].copy()


# ---------------------------------------------------------------- device ----

def neighbours(xyz, rc, stream=None):
    """Cell-list neighbour CSR of device float64 coordinates [n,3] -> (indptr, indices)."""
    require_gpu()
    _check_dev(xyz)
    if xyz.dtype != torch.float64 or xyz.dim() != 2 or xyz.shape[1] != 3:
        raise ValueError("xyz must be float64 [n,3]")
    n = xyz.shape[0]
    indptr = torch.empty(n + 1, dtype=torch.int64, device=xyz.device)
    nnz = C.c_int64(0)
    call("inch_neighbors_celllist", n, P(xyz), float(rc), P(indptr), None, 0, C.byref(nnz), stream_ptr(stream))
    indices = torch.empty(max(1, nnz.value), dtype=torch.int32, device=xyz.device)
    call("inch_neighbors_celllist", n, P(xyz), float(rc), P(indptr), P(indices), nnz.value, C.byref(nnz),
         stream_ptr(stream))
    return indptr, indices[:nnz.value]


def csr_permute(indptr, indices, order, undo, stream=None):
    """The neighbour CSR renumbered by a permutation (inch_csr_permute): order int32 [n]
    (new -> old), undo int64 [n] (old -> new), device tensors. Returns (indptr, indices) of
    the reordered atoms, equal to a fresh neighbour search of the reordered coordinates."""
    require_gpu()
    n = indptr.numel() - 1
    if (order.dtype != torch.int32 or undo.dtype != torch.int64 or order.numel() != n or undo.numel() != n
            or indptr.dtype != torch.int64 or indices.dtype != torch.int32):
        raise ValueError("csr_permute: indptr int64 [n+1], indices int32, order int32 [n], undo int64 [n]")
    out_ptr = torch.empty(n + 1, dtype=torch.int64, device=indptr.device)
    out_idx = torch.empty(max(1, indices.numel()), dtype=torch.int32, device=indptr.device)
    call("inch_csr_permute", n, P(indptr), P(indices), P(order), P(undo), P(out_ptr), P(out_idx), stream_ptr(stream))
    return out_ptr, out_idx[:indices.numel()]


def neighbours_pbc(xyz, rc, box, origin, stream=None):
    """Neighbour CSR with periodic images in x and y: (indptr, indices, images), images[e]
    the 9-bit mask of the images of column j within the cutoff (see inch_neighbors_celllist_pbc).
    box = (box_x, box_y), origin = (x0, y0) of the periodic box."""
    require_gpu()
    _check_dev(xyz)
    if xyz.dtype != torch.float64 or xyz.dim() != 2 or xyz.shape[1] != 3:
        raise ValueError("xyz must be float64 [n,3]")
    n = xyz.shape[0]
    bx, by = float(box[0]), float(box[1])
    ox, oy = float(origin[0]), float(origin[1])
    indptr = torch.empty(n + 1, dtype=torch.int64, device=xyz.device)
    nnz = C.c_int64(0)
    call("inch_neighbors_celllist_pbc", n, P(xyz), float(rc), bx, by, ox, oy, P(indptr), None, None, 0,
         C.byref(nnz), stream_ptr(stream))
    indices = torch.empty(max(1, nnz.value), dtype=torch.int32, device=xyz.device)
    images = torch.empty(max(1, nnz.value), dtype=torch.uint16, device=xyz.device)
    call("inch_neighbors_celllist_pbc", n, P(xyz), float(rc), bx, by, ox, oy, P(indptr), P(indices), P(images),
         nnz.value, C.byref(nnz), stream_ptr(stream))
    return indptr, indices[:nnz.value], images[:nnz.value]


def hessian_assemble_pbc(xyz, plus_i, rowptr, col, images, box, want64=True, want32=True, stream=None):
    require_gpu()
    _check_dev(xyz, rowptr, col)
    n = xyz.shape[0]
    nnzb = col.shape[0]
    if rowptr.shape[0] != n + 1 or images.shape[0] != nnzb:
        raise ValueError("rowptr / images inconsistent with the pattern")
    v64 = torch.empty(nnzb * 9, dtype=torch.float64, device=xyz.device) if want64 else None
    v32 = torch.empty(nnzb * 9, dtype=torch.float32, device=xyz.device) if want32 else None
    diag = torch.empty(3 * n, dtype=torch.float64, device=xyz.device)
    call("inch_hessian_assemble_pbc", n, P(xyz), float(plus_i), P(rowptr), P(col), P(images), float(box[0]),
         float(box[1]), P(v64), P(v32), P(diag), stream_ptr(stream))
    return v64, v32, diag


def matfree_pbc(xyz, rowptr, col, images, box, plus_i, x, y, s1=1.0, c0=0.0, z=None, c2=0.0, acc=None, mu=0.0,
                Q=None, qcoef=None, stream=None):
    """Matrix-free ANM product with periodic images; x, y [3n] (fused epilogue) or [3n, nrhs]."""
    n = xyz.shape[0]
    nd = 3 * n
    nrhs = 1 if x.dim() == 1 else x.shape[1]
    if x.shape[0] != nd or y.shape != x.shape:
        raise ValueError("x, y must be [3n] or [3n, nrhs]")
    vd = dt(x)
    for t in (y, z, acc):
        if t is not None and dt(t) != vd:
            raise ValueError("all vectors must share a dtype")
    mq = 0
    if qcoef is not None:
        if Q is None or Q.dtype != torch.float64 or Q.shape[1] != nd:
            raise ValueError("Q must be float64 [m, 3n]")
        mq = Q.shape[0]
    call("inch_anm_matfree_pbc", n, P(xyz), P(rowptr), P(col), P(images), float(box[0]), float(box[1]),
         float(plus_i), vd, P(x), P(y), nrhs, float(s1), float(c0), P(z), float(c2), P(acc), float(mu), P(Q),
         P(qcoef), mq, stream_ptr(stream))


def hessian_assemble(xyz, rc, plus_i, rowptr, col, want64=True, want32=True, stream=None):
    require_gpu()
    _check_dev(xyz, rowptr, col)
    n = xyz.shape[0]
    nnzb = col.shape[0]
    if rowptr.shape[0] != n + 1:
        raise ValueError("rowptr must have n+1 entries")
    v64 = torch.empty(nnzb * 9, dtype=torch.float64, device=xyz.device) if want64 else None
    v32 = torch.empty(nnzb * 9, dtype=torch.float32, device=xyz.device) if want32 else None
    diag = torch.empty(3 * n, dtype=torch.float64, device=xyz.device)
    call("inch_hessian_assemble", n, P(xyz), float(rc), float(plus_i), P(rowptr), P(col), P(v64), P(v32),
         P(diag), stream_ptr(stream))
    return v64, v32, diag


def bsr_col16(n_atoms, rowptr, col):
    """(rowbase int32 [n], col16 uint16 [nnzb]): 16-bit column offsets from each row's first
    column (raises InchingError when a row spans 65,536 columns or more)."""
    rowbase = torch.empty(n_atoms, dtype=torch.int32, device=col.device)
    col16 = torch.empty(max(1, col.numel()), dtype=torch.uint16, device=col.device)
    call("inch_bsr_col16", n_atoms, P(rowptr), P(col), P(rowbase), P(col16), stream_ptr(None))
    return rowbase, col16[:col.numel()]


CLUSTER_ATOMS = 4   # atoms per cluster of the cluster layout (inch_anm_clusters)


def anm_clusters(xyz, rowptr, col, cell_xy, cell_z, stream=None, one_pass=True):
    """Cluster layout of the matrix-free operator (inch_anm_clusters_one_pass, or the two calls
    of inch_anm_clusters with ``one_pass=False``): (cluster_atoms int32 [ceil(n/4)*4],
    cl_ptr int64 [ncl+1], base int32 [ncl], entries int32 [nnz] holding the uint32 words
    ((j - base) << 8) | mask). Raises InchingError (INCH_ERR_CAPACITY) when the layout does
    not fit the matrix."""
    require_gpu()
    _check_dev(xyz)
    n = xyz.shape[0]
    if rowptr.numel() != n + 1:
        raise ValueError("rowptr must have n + 1 entries")
    ncl = (n + CLUSTER_ATOMS - 1) // CLUSTER_ATOMS
    dev = xyz.device
    atoms = torch.empty(ncl * CLUSTER_ATOMS, dtype=torch.int32, device=dev)
    ptr = torch.empty(ncl + 1, dtype=torch.int64, device=dev)
    base = torch.empty(ncl, dtype=torch.int32, device=dev)
    nnz = C.c_int64(0)
    if one_pass:
        room = max(1, int(rowptr[-1].item()))          # the CSR's entries bound the layout's
        entries = torch.empty(room, dtype=torch.int32, device=dev)
        call("inch_anm_clusters_one_pass", n, P(xyz), P(rowptr), P(col), float(cell_xy), float(cell_z), P(atoms),
             P(ptr), P(base), P(entries), room, C.byref(nnz), stream_ptr(stream))
        return atoms, ptr, base, entries[:nnz.value].clone()
    call("inch_anm_clusters", n, P(xyz), P(rowptr), P(col), float(cell_xy), float(cell_z), P(atoms), P(ptr), P(base),
         None, 0, C.byref(nnz), stream_ptr(stream))
    entries = torch.empty(max(1, nnz.value), dtype=torch.int32, device=dev)
    call("inch_anm_clusters", n, P(xyz), P(rowptr), P(col), float(cell_xy), float(cell_z), P(atoms), P(ptr), P(base),
         P(entries), nnz.value, C.byref(nnz), stream_ptr(stream))
    return atoms, ptr, base, entries[:nnz.value]


def anm_records(xyz, clusters, copies=3, stream=None):
    """Record form of the cluster layout (inch_anm_records): (list of ``copies`` record
    buffers float64 [24 ncl], each with the slots' coordinates and a zero vector part,
    rec_entries int32 [nnz] holding the uint32 words 48 slot | mask: the byte offset of the
    neighbour's 48-byte record with the mask in its low 4 bits)."""
    require_gpu()
    _check_dev(xyz)
    n = xyz.shape[0]
    atoms, ptr, base, entries = clusters
    ncl = (n + CLUSTER_ATOMS - 1) // CLUSTER_ATOMS
    if atoms.numel() != CLUSTER_ATOMS * ncl or ptr.numel() != ncl + 1 or base.numel() != ncl:
        raise ValueError("cluster layout does not match xyz")
    rec = torch.empty(24 * ncl, dtype=torch.float64, device=xyz.device)
    rent = torch.empty(max(1, entries.numel()), dtype=torch.int32, device=xyz.device)
    call("inch_anm_records", n, P(xyz), ncl, P(atoms), P(ptr), P(base), P(entries), P(rec), P(rent),
         stream_ptr(stream))
    return [rec] + [rec.clone() for _ in range(copies - 1)], rent[:entries.numel()]


def matfree_rec(rec_in, rec_z, rec_out, cl_ptr, rec_entries, plus_i, s1=1.0, c0=0.0, c2=0.0, acc=None, mu=0.0,
                xcd=True, y=None, stream=None):
    """One Chebyshev step on records (inch_anm_matfree_rec); ``y``: optional slot-order output."""
    ncl = cl_ptr.numel() - 1
    for r in (rec_in, rec_z, rec_out):
        if r is not None and (r.numel() != 24 * ncl or r.dtype != torch.float64):
            raise ValueError("records must be float64 [24 ncl]")
    for t in (acc, y):
        if t is not None and (t.numel() != 12 * ncl or t.dtype != torch.float64):
            raise ValueError("acc and y must be float64 [12 ncl]")
    call("inch_anm_matfree_rec", ncl, P(rec_in), P(rec_z), P(rec_out), P(y), P(cl_ptr), P(rec_entries),
         float(plus_i), float(s1), float(c0), float(c2), P(acc), float(mu), 1 if xcd else 0, stream_ptr(stream))


def rec_pack(cluster_atoms, v, rec, acc=None, mu0=0.0, stream=None, ncl=None):
    """Vector part of the records from v (matrix order through cluster_atoms, or slot order
    when cluster_atoms is None, then ``ncl`` is required)."""
    if cluster_atoms is not None:
        ncl = cluster_atoms.numel() // CLUSTER_ATOMS
    call("inch_rec_pack", int(ncl), P(cluster_atoms), P(v), P(rec), P(acc), float(mu0), stream_ptr(stream))


def rec_unpack(cluster_atoms, acc, u, stream=None):
    ncl = cluster_atoms.numel() // CLUSTER_ATOMS
    call("inch_rec_unpack", ncl, P(cluster_atoms), P(acc), P(u), stream_ptr(stream))


def matfree_cl_apply(xyz, clusters, plus_i, X, Y, stream=None):
    """Y = (H + plus_i I) X on the cluster layout, float64 X [3n] or [3n, r] row-major."""
    n = xyz.shape[0]
    X2 = X.reshape(3 * n, -1)
    Y2 = Y.reshape(3 * n, -1)
    if X2.dtype != torch.float64 or Y2.dtype != torch.float64 or not X2.is_contiguous() or not Y2.is_contiguous():
        raise ValueError("X and Y must be contiguous float64 [3n, r]")
    if Y2.shape != X2.shape:
        raise ValueError("X and Y shapes differ")
    atoms, ptr, base, entries = clusters
    ncl = (n + CLUSTER_ATOMS - 1) // CLUSTER_ATOMS
    if atoms.numel() != CLUSTER_ATOMS * ncl or ptr.numel() != ncl + 1 or base.numel() != ncl:
        raise ValueError("cluster layout does not match n")
    call("inch_anm_matfree_cl_apply", n, P(xyz), ncl, P(atoms), P(ptr), P(base), P(entries), float(plus_i), P(X2),
         P(Y2), X2.shape[1], stream_ptr(stream))


def matfree_cl(xyz, clusters, plus_i, x, y, s1=1.0, c0=0.0, z=None, c2=0.0, acc=None, mu=0.0, Q=None, qcoef=None,
               config=0, stream=None):
    """Matrix-free ANM product with the fused epilogue on the cluster layout ``clusters``
    (the tuple of anm_clusters)."""
    n = xyz.shape[0]
    nd = 3 * n
    for t in (x, y, z, acc):
        if t is not None and t.numel() != nd:
            raise ValueError("vector length mismatch")
    vd = dt(x)
    for t in (y, z, acc):
        if t is not None and dt(t) != vd:
            raise ValueError("all vectors must share a dtype")
    mq = 0
    if qcoef is not None:
        if Q is None or Q.dtype != torch.float64 or Q.shape[1] != nd:
            raise ValueError("Q must be float64 [m, 3n]")
        mq = Q.shape[0]
    atoms, ptr, base, entries = clusters
    ncl = (n + CLUSTER_ATOMS - 1) // CLUSTER_ATOMS
    if atoms.numel() != CLUSTER_ATOMS * ncl or ptr.numel() != ncl + 1 or base.numel() != ncl:
        raise ValueError("cluster layout does not match n")
    call("inch_anm_matfree_cl", n, P(xyz), ncl, P(atoms), P(ptr), P(base), P(entries), float(plus_i), vd, P(x), P(y),
         float(s1), float(c0), P(z), float(c2), P(acc), float(mu), P(Q), P(qcoef), mq, int(config), stream_ptr(stream))


def bsr_spmv(n_atoms, rowptr, col, vals, x, y, s1=1.0, c0=0.0, z=None, c2=0.0, acc=None, mu=0.0, Q=None,
             qcoef=None, lanes=16, stream=None, col16=None):
    """y = s1 (A x - c0 x) - c2 z (+ sum_m qcoef[m] Q[m]); acc += mu y (all terms optional).
    ``col16`` = (rowbase, col16) selects the 16-bit column-offset kernel (col is then unused)."""
    nd = 3 * n_atoms
    for t in (x, y, z, acc):
        if t is not None and t.numel() != nd:
            raise ValueError("vector length mismatch")
    if vals.numel() != 9 * col.numel() or rowptr.numel() != n_atoms + 1:
        raise ValueError("matrix arrays inconsistent")
    vd = dt(x)
    for t in (y, z, acc):
        if t is not None and dt(t) != vd:
            raise ValueError("all vectors must share a dtype")
    mq = 0
    if qcoef is not None:
        if Q is None or Q.dtype != torch.float64 or Q.shape[1] != nd:
            raise ValueError("Q must be float64 [m, 3n]")
        mq = Q.shape[0]
        if qcoef.numel() < mq:
            raise ValueError("qcoef too short")
    if col16 is not None:
        call("inch_bsr_spmv16", n_atoms, P(rowptr), P(col16[0]), P(col16[1]), P(vals), dt(vals), vd, P(x), P(y),
             float(s1), float(c0), P(z), float(c2), P(acc), float(mu), P(Q), P(qcoef), mq, int(lanes),
             stream_ptr(stream))
        return
    call("inch_bsr_spmv", n_atoms, P(rowptr), P(col), P(vals), dt(vals), vd, P(x), P(y), float(s1), float(c0),
         P(z), float(c2), P(acc), float(mu), P(Q), P(qcoef), mq, int(lanes), stream_ptr(stream))


def bsr_spmv_part(n_atoms, rowptr, col, vals, x, y, part, stream=None):
    """part 0: y = A x, 1: y = L x, 2: y = L^T x (reference half-storage views)."""
    if x.numel() != 3 * n_atoms or y.numel() != 3 * n_atoms:
        raise ValueError("vector length mismatch")
    call("inch_bsr_spmv_part", n_atoms, P(rowptr), P(col), P(vals), dt(vals), dt(x), P(x), P(y), int(part),
         stream_ptr(stream))


def bsr_spmm(n_atoms, rowptr, col, vals, X, Y, stream=None):
    nd = 3 * n_atoms
    if X.shape != Y.shape or X.shape[0] != nd or X.dim() != 2:
        raise ValueError("X, Y must be [3n, nrhs]")
    call("inch_bsr_spmm", n_atoms, P(rowptr), P(col), P(vals), dt(vals), dt(X), P(X), P(Y), X.shape[1],
         stream_ptr(stream))


def matfree_apply(xyz, rowptr, col, plus_i, X, Y, stream=None):
    n = xyz.shape[0]
    if X.dim() == 1:
        X2, Y2 = X.view(-1, 1), Y.view(-1, 1)
    else:
        X2, Y2 = X, Y
    if X2.shape[0] != 3 * n or X2.shape != Y2.shape or X2.dtype != torch.float64 or Y2.dtype != torch.float64:
        raise ValueError("X, Y must be float64 [3n, nrhs]")
    call("inch_anm_matfree_apply", n, P(xyz), P(rowptr), P(col), float(plus_i), P(X2), P(Y2), X2.shape[1],
         stream_ptr(stream))


def matfree_spmv(xyz, rowptr, col, plus_i, x, y, s1=1.0, c0=0.0, z=None, c2=0.0, acc=None, mu=0.0, Q=None,
                 qcoef=None, config=0, stream=None, col16=None):
    """Matrix-free ANM product with the fused epilogue (see bsr_spmv); ``col16`` = (rowbase,
    col16) selects 16-bit column offsets."""
    n = xyz.shape[0]
    nd = 3 * n
    for t in (x, y, z, acc):
        if t is not None and t.numel() != nd:
            raise ValueError("vector length mismatch")
    vd = dt(x)
    for t in (y, z, acc):
        if t is not None and dt(t) != vd:
            raise ValueError("all vectors must share a dtype")
    mq = 0
    if qcoef is not None:
        if Q is None or Q.dtype != torch.float64 or Q.shape[1] != nd:
            raise ValueError("Q must be float64 [m, 3n]")
        mq = Q.shape[0]
    if col16 is not None:
        call("inch_anm_matfree_spmv16", n, P(xyz), P(rowptr), P(col16[0]), P(col16[1]), float(plus_i), vd, P(x), P(y),
             float(s1), float(c0), P(z), float(c2), P(acc), float(mu), P(Q), P(qcoef), mq, int(config),
             stream_ptr(stream))
        return
    call("inch_anm_matfree_spmv", n, P(xyz), P(rowptr), P(col), float(plus_i), vd, P(x), P(y), float(s1),
         float(c0), P(z), float(c2), P(acc), float(mu), P(Q), P(qcoef), mq, int(config), stream_ptr(stream))


def stream_copy(src, dst, stream=None):
    nb = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() < nb:
        raise ValueError("dst too small")
    call("inch_stream_copy", nb, P(src), P(dst), stream_ptr(stream))


class Workspace:
    """Scratch for the two-phase reductions, sized for up to ``J`` dot products."""

    def __init__(self, J, device):
        self.J = max(1, int(J))
        from ._lib import lib
        self.buf = torch.empty(int(lib().inch_multidot_work_size(self.J)), dtype=torch.float64, device=device)


def multidot(V, J, u, h, ws, stream=None):
    """h[:J] = V[:J] @ u  (V [rows, len] float64 row-major)."""
    if J > ws.J:
        raise ValueError("workspace too small")
    if V.shape[0] < J or V.shape[1] != u.numel() or h.numel() < J:
        raise ValueError("shape mismatch in multidot")
    call("inch_multidot", u.numel(), int(J), P(V), V.stride(0), P(u), P(h), P(ws.buf), stream_ptr(stream))


def gemv_update(V, J, h, u, scale=1.0, nrm2_out=None, ws=None, stream=None):
    """u -= scale * V[:J]^T h[:J]; optionally nrm2_out[0] = ||u||."""
    if J > 0 and (V.shape[0] < J or V.shape[1] != u.numel() or h.numel() < J):
        raise ValueError("shape mismatch in gemv_update")
    call("inch_gemv_update", u.numel(), int(J), P(V) if J > 0 else None, V.stride(0) if J > 0 else u.numel(),
         P(h) if J > 0 else None, float(scale), P(u), P(nrm2_out), P(ws.buf) if ws is not None else None,
         stream_ptr(stream))


def scale_div(src, div, dst, dst2=None, stream=None):
    call("inch_vec_scale_div", src.numel(), P(src), P(div), P(dst), P(dst2), stream_ptr(stream))


def axpy_dev(a, sign, x, y, stream=None):
    call("inch_vec_axpy_dev", x.numel(), P(a), float(sign), P(x), P(y), stream_ptr(stream))


def axpby(a, x, b, y, stream=None):
    if x.numel() != y.numel():
        raise ValueError("length mismatch")
    call("inch_vec_axpby", x.numel(), float(a), P(x), dt(x), float(b), P(y), dt(y), stream_ptr(stream))


def convert(src, dst, stream=None):
    if src.numel() != dst.numel():
        raise ValueError("length mismatch")
    call("inch_vec_convert", src.numel(), P(src), dt(src), P(dst), dt(dst), stream_ptr(stream))


def rr_combine(V, m, W, out, stream=None):
    """out[:k] = W^T V[:m]   (W float64 [m, k] on device, out [>=k, len])."""
    k = W.shape[1]
    if W.shape[0] != m or V.shape[0] < m or out.shape[0] < k or out.shape[1] != V.shape[1]:
        raise ValueError("shape mismatch in rr_combine")
    if out.data_ptr() == V.data_ptr():
        raise ValueError("rr_combine output must not alias V")
    call("inch_rr_combine", V.shape[1], int(m), int(k), P(V), V.stride(0), P(W.contiguous()), P(out),
         out.stride(0), stream_ptr(stream))


class GramWorkspace:
    def __init__(self, k1, k2, device):
        from ._lib import lib
        self.k1, self.k2 = int(k1), int(k2)
        size = max(int(lib().inch_gram_work_size(self.k1, self.k2)), 256 * max(self.k1, self.k2))
        self.buf = torch.empty(size, dtype=torch.float64, device=device)


def gram(X, Y, G=None, ws=None, stream=None):
    """G = X Y^T for float64 row blocks X [k1, n], Y [k2, n] (MFMA kernel)."""
    k1, k2 = X.shape[0], Y.shape[0]
    if X.shape[1] != Y.shape[1] or X.dtype != torch.float64 or Y.dtype != torch.float64:
        raise ValueError("gram: X and Y must be float64 with equal row length")
    if X.stride(1) != 1 or Y.stride(1) != 1:
        raise ValueError("gram: rows must be contiguous")
    if ws is None or ws.k1 < k1 or ws.k2 < k2 or ws.k1 * ws.k2 < k1 * k2:
        ws = GramWorkspace(k1, k2, X.device)
    if G is None:
        G = torch.empty((k1, k2), dtype=torch.float64, device=X.device)
    call("inch_gram", X.shape[1], k1, P(X), X.stride(0), k2, P(Y), Y.stride(0), P(G), P(ws.buf), stream_ptr(stream))
    return G


def residual_norms(X, Y, theta, out=None, ws=None, stream=None):
    """out[j] = ||Y[j] - theta[j] X[j]|| for float64 row blocks."""
    k = X.shape[0]
    if Y.shape != X.shape or theta.numel() < k:
        raise ValueError("residual_norms: shape mismatch")
    if ws is None or ws.buf.numel() < 256 * k:
        ws = GramWorkspace(k, 1, X.device)
    if out is None:
        out = torch.empty(k, dtype=torch.float64, device=X.device)
    call("inch_residual_norms", X.shape[1], k, P(X), X.stride(0), P(Y), Y.stride(0), P(theta), P(out), P(ws.buf),
         stream_ptr(stream))
    return out


CGS2_MAX_J = 256


class CGS2Workspace:
    def __init__(self, J, device):
        from ._lib import lib
        self.J = int(J)
        self.buf = torch.empty(int(lib().inch_cgs2_work_size(self.J)), dtype=torch.float64, device=device)


class DCGS2Workspace:
    def __init__(self, J, device):
        from ._lib import lib
        self.J = int(J)
        self.buf = torch.empty(int(lib().inch_dcgs2_work_size(self.J)), dtype=torch.float64, device=device)


def dcgs2_step(V, J, u, alpha, beta, beta_out, vnext=None, ws=None, stream=None):
    """One Lanczos orthogonalisation step with delayed reorthogonalisation (inch_dcgs2_step):
    i = J - 1, u = Op V[i] for the provisional V[i]; corrects V[i] against V[:i], projects u
    off the corrected V[:J], writes alpha[i] (using beta[i - 1]), beta_out = |u| and
    vnext = u / |u|."""
    if J < 1:
        raise ValueError("dcgs2_step: J must be >= 1")
    if V.shape[0] < J or V.shape[1] != u.numel() or alpha.numel() < J or (J > 1 and beta.numel() < J - 1):
        raise ValueError("dcgs2_step: shape mismatch")
    if ws is None or ws.J < J:
        ws = DCGS2Workspace(J, u.device)
    call("inch_dcgs2_step", u.numel(), int(J), P(V), V.stride(0), P(u), P(alpha), P(beta), P(beta_out), P(vnext),
         P(ws.buf), stream_ptr(stream))


def cgs2_reorth(V, J, h1, u, beta_out, vnext=None, ws=None, stream=None):
    """u' = u - V[:J]^T h1; beta = ||u'||; u = u' - V[:J]^T (V[:J] u'); vnext = u / beta."""
    if J > CGS2_MAX_J or J < 1:
        raise ValueError("cgs2_reorth: J must be in 1..256")
    if V.shape[0] < J or V.shape[1] != u.numel() or h1.numel() < J:
        raise ValueError("cgs2_reorth: shape mismatch")
    if ws is None or ws.J < J:
        ws = CGS2Workspace(J, u.device)
    call("inch_cgs2_reorth", u.numel(), int(J), P(V), V.stride(0), P(h1), P(u), P(vnext), P(beta_out), P(ws.buf),
         stream_ptr(stream))
