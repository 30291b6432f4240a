"""ctypes binding of libinching_hip.so (the C ABI declared in include/inching.h).

The library is required: importing this module without the built library raises
(there is no CPU fallback anywhere in the product path).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libinching_hip.so")

F32, F64 = 0, 1

_i32, _i64, _u64, _dbl, _vp = C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_void_p
_STATUS = C.c_int32

# name -> (restype, argtypes)
SIGNATURES = {
    "inch_status_string": (C.c_char_p, [_i32]),
    "inch_last_error": (C.c_char_p, []),
    "inch_abi_version": (_i32, []),
    "inch_synth_globule": (_STATUS, [_i64, _dbl, _dbl, _u64, _i64, _vp]),
    "inch_rcm_host": (_STATUS, [_i64, _vp, _vp, _vp, _i32, _vp, _vp]),
    "inch_batch_head": (_STATUS, [_i64, _i64, _vp, _i64, _vp]),
    "inch_neighbors_celllist": (_STATUS, [_i64, _vp, _dbl, _vp, _vp, _i64, _vp, _vp]),
    "inch_csr_permute": (_STATUS, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_hessian_assemble": (_STATUS, [_i64, _vp, _dbl, _dbl, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_bsr_spmv": (_STATUS, [_i64, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _dbl, _dbl, _vp, _dbl, _vp, _dbl,
                                _vp, _vp, _i32, _i32, _vp]),
    "inch_bsr_spmv_part": (_STATUS, [_i64, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _i32, _vp]),
    "inch_bsr_spmm": (_STATUS, [_i64, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _i32, _vp]),
    "inch_anm_matfree_apply": (_STATUS, [_i64, _vp, _vp, _vp, _dbl, _vp, _vp, _i32, _vp]),
    "inch_anm_matfree_spmv": (_STATUS, [_i64, _vp, _vp, _vp, _dbl, _i32, _vp, _vp, _dbl, _dbl, _vp, _dbl, _vp,
                                        _dbl, _vp, _vp, _i32, _i32, _vp]),
    "inch_rcm_device_work_size": (_i64, [_i64]),
    "inch_rcm_device": (_STATUS, [_i64, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _i64, _vp, _i64, _vp, _vp]),
    "inch_neighbors_celllist_pbc": (_STATUS, [_i64, _vp, _dbl, _dbl, _dbl, _dbl, _dbl, _vp, _vp, _vp, _i64, _vp,
                                              _vp]),
    "inch_hessian_assemble_pbc": (_STATUS, [_i64, _vp, _dbl, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp, _vp]),
    "inch_anm_matfree_pbc": (_STATUS, [_i64, _vp, _vp, _vp, _vp, _dbl, _dbl, _dbl, _i32, _vp, _vp, _i32, _dbl, _dbl,
                                       _vp, _dbl, _vp, _dbl, _vp, _vp, _i32, _vp]),
    "inch_anm_matfree_spmv16": (_STATUS, [_i64, _vp, _vp, _vp, _vp, _dbl, _i32, _vp, _vp, _dbl, _dbl, _vp, _dbl, _vp,
                                          _dbl, _vp, _vp, _i32, _i32, _vp]),
    "inch_bsr_spmv16": (_STATUS, [_i64, _vp, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _dbl, _dbl, _vp, _dbl, _vp, _dbl,
                                  _vp, _vp, _i32, _i32, _vp]),
    "inch_bsr_col16": (_STATUS, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "inch_stream_copy": (_STATUS, [_i64, _vp, _vp, _vp]),
    "inch_anm_clusters": (_STATUS, [_i64, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "inch_anm_clusters_one_pass": (_STATUS, [_i64, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "inch_anm_matfree_cl_apply": (_STATUS, [_i64, _vp, _i64, _vp, _vp, _vp, _vp, _dbl, _vp, _vp, _i32, _vp]),
    "inch_anm_records": (_STATUS, [_i64, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_anm_matfree_rec": (_STATUS, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _dbl, _dbl, _vp, _dbl, _i32,
                                       _vp]),
    "inch_rec_pack": (_STATUS, [_i64, _vp, _vp, _vp, _vp, _dbl, _vp]),
    "inch_rec_unpack": (_STATUS, [_i64, _vp, _vp, _vp, _vp]),
    "inch_anm_matfree_cl": (_STATUS, [_i64, _vp, _i64, _vp, _vp, _vp, _vp, _dbl, _i32, _vp, _vp, _dbl, _dbl, _vp,
                                      _dbl, _vp, _dbl, _vp, _vp, _i32, _i32, _vp]),
    "inch_multidot_work_size": (_i64, [_i32]),
    "inch_multidot": (_STATUS, [_i64, _i32, _vp, _i64, _vp, _vp, _vp, _vp]),
    "inch_gemv_update": (_STATUS, [_i64, _i32, _vp, _i64, _vp, _dbl, _vp, _vp, _vp, _vp]),
    "inch_vec_scale_div": (_STATUS, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "inch_vec_axpy_dev": (_STATUS, [_i64, _vp, _dbl, _vp, _vp, _vp]),
    "inch_vec_convert": (_STATUS, [_i64, _vp, _i32, _vp, _i32, _vp]),
    "inch_vec_axpby": (_STATUS, [_i64, _dbl, _vp, _i32, _dbl, _vp, _i32, _vp]),
    "inch_gram_work_size": (_i64, [_i32, _i32]),
    "inch_gram": (_STATUS, [_i64, _i32, _vp, _i64, _i32, _vp, _i64, _vp, _vp, _vp]),
    "inch_residual_norms": (_STATUS, [_i64, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp]),
    "inch_cgs2_work_size": (_i64, [_i32]),
    "inch_cgs2_reorth": (_STATUS, [_i64, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_dcgs2_work_size": (_i64, [_i32]),
    "inch_dcgs2_step": (_STATUS, [_i64, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_dcgs2_step_noalloc": (_STATUS, [_i64, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_seg_multidot": (_STATUS, [_i32, _i32, _vp, _vp, _vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp]),
    "inch_seg_update": (_STATUS, [_i32, _i32, _vp, _vp, _vp, _vp, _i32, _vp, _i64, _vp, _i32, _dbl, _vp, _vp, _vp,
                                  _vp]),
    "inch_seg_scale": (_STATUS, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_seg_rigid_apply": (_STATUS, [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _vp, _vp,
                                       _vp]),
    "inch_seg_cgs_b": (_STATUS, [_i32, _i32, _vp, _vp, _vp, _vp, _i32, _vp, _i64, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "inch_seg_dcgs2_work_size": (_i64, [_i32, _i32, _i32]),
    "inch_seg_dcgs2_step": (_STATUS, [_i32, _i32, _vp, _vp, _vp, _vp, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_seg_combine": (_STATUS, [_i32, _vp, _vp, _vp, _i32, _i32, _vp, _i64, _vp, _vp, _i64, _vp]),
    "inch_seg_resnorm": (_STATUS, [_i32, _i32, _vp, _vp, _vp, _vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _vp,
                                   _vp]),
    "inch_gnn_dense_fwd": (_STATUS, [_i64, _i32, _i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "inch_gnn_work_size": (_i64, [_i32, _i32]),
    "inch_gnn_dense_bwd": (_STATUS, [_i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_gnn_aggregate": (_STATUS, [_i64, _i32, _vp, _vp, _vp, _i32, _vp, _i32, _i32, _vp]),
    "inch_gnn_relu_bwd": (_STATUS, [_i64, _vp, _vp, _vp, _vp]),
    "inch_rr_combine": (_STATUS, [_i64, _i32, _i32, _vp, _i64, _vp, _vp, _i64, _vp]),
    "inch_sqfluct": (_STATUS, [_i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "inch_cross_correlation": (_STATUS, [_i32, _vp, _i32, _i64, _vp, _vp, _i32, _vp, _vp]),
}


class InchingError(RuntimeError):
    def __init__(self, name, status, msg):
        super().__init__("%s failed (%d): %s" % (name, status, msg))
        self.status = status


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libinching_hip.so is not built (expected at %s); run "
                              "`python -m inching_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


_fns = {}


def call(name, *args):
    """Call an ``inch_*`` entry point and raise InchingError on a non-zero status."""
    fn = _fns.get(name)
    if fn is None:
        fn = _fns[name] = getattr(lib(), name)
    st = fn(*args)
    if SIGNATURES[name][0] is _STATUS and st != 0:
        msg = lib().inch_last_error()
        raise InchingError(name, st, msg.decode() if msg else "")
    return st


def exported_symbols():
    return sorted(SIGNATURES)
