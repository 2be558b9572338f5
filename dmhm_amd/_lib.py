"""ctypes binding of libdmhm_amd.so (the C-ABI in include/dmhm_amd.h).

The library must have been built (``python -c 'import __graft_entry__ as g;
g.build()'`` or ``make -C dmhm_amd``); importing this module without it raises
immediately: there is no CPU fallback for any operation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdmhm_amd.so")

DMHM_OK = 0
DMHM_F64 = 8
DMHM_F32 = 4
DMHM_HOST = 0
DMHM_DEVICE = 1
DMHM_RANK_FEATURES = 12

_ERRORS = {1: "DMHM_EINVAL", 2: "DMHM_ENONCONF", 3: "DMHM_ESINGULAR", 4: "DMHM_ENOCONV",
           5: "DMHM_EHIP", 6: "DMHM_ENOMEM", 7: "DMHM_ENODEV"}


class DmhmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_ERRORS.get(code, str(code)), msg))
        self.code = code


class NonConformalError(DmhmError, ValueError):
    pass


class NotConvergedError(DmhmError):
    pass


class Info(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n", ctypes.c_int), ("numLevels", ctypes.c_int),
                ("maxRank", ctypes.c_int), ("strong", ctypes.c_int), ("precision", ctypes.c_int),
                ("numDense", ctypes.c_longlong), ("numLowRank", ctypes.c_longlong),
                ("storageBytes", ctypes.c_longlong), ("sumRank", ctypes.c_longlong),
                ("maxBlockRank", ctypes.c_int)]


class HHOpts(ctypes.Structure):
    _fields_ = [("oversample", ctypes.c_int), ("tolMode", ctypes.c_int), ("relTol", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("rankModel", ctypes.c_void_p), ("rankMargin", ctypes.c_double)]


class SchulzOpts(ctypes.Structure):
    _fields_ = [("numIterations", ctypes.c_int), ("maxIterations", ctypes.c_int),
                ("theta", ctypes.c_double), ("confidence", ctypes.c_double),
                ("stopTol", ctypes.c_double), ("mixed", ctypes.c_int), ("switchTol", ctypes.c_double),
                ("hh", HHOpts), ("stallStop", ctypes.c_int), ("graph", ctypes.c_int), ("bf16", ctypes.c_int),
                ("bf16Tol", ctypes.c_double), ("start", ctypes.c_int)]


class SchulzReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("iterationsLowPrecision", ctypes.c_int),
                ("residualEstimate", ctypes.c_double), ("normEstimate", ctypes.c_double),
                ("seconds", ctypes.c_double), ("history", ctypes.c_double * 128),
                ("graphIterations", ctypes.c_int), ("iterationsBf16", ctypes.c_int)]


_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_llp = ctypes.POINTER(ctypes.c_longlong)
_fp = ctypes.POINTER(ctypes.c_float)
_h = ctypes.c_void_p
_hp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); every symbol declared in include/dmhm_amd.h.
SIGNATURES = {
    "dmhm_init": (ctypes.c_int, [ctypes.c_int]),
    "dmhm_finalize": (ctypes.c_int, []),
    "dmhm_last_error": (ctypes.c_char_p, []),
    "dmhm_synchronize": (ctypes.c_int, []),
    "dmhm_build_natural_to_hierarchical_map": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _ip]),
    "dmhm_build_natural_to_hierarchical_map_device": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, _ip]),
    "dmhm_geometry_block_counts": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _llp, _llp,
                                                  ctypes.c_int]),
    "dmhm_geometry_leaves": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int,
                                            _ip]),
    "dmhm_geometry_product_terms": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _llp]),
    "dmhm_hmat_create_from_csr": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, _ip, _ip, _dp, _hp]),
    "dmhm_hmat_create_zero": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, _hp]),
    "dmhm_hmat_destroy": (ctypes.c_int, [_h]),
    "dmhm_hmat_info": (ctypes.c_int, [_h, ctypes.POINTER(Info)]),
    "dmhm_hmat_block_counts": (ctypes.c_int, [_h, _llp, _llp, ctypes.c_int]),
    "dmhm_hmat_set_to_random": (ctypes.c_int, [_h, ctypes.c_uint64]),
    "dmhm_hmat_multiply_dense": (ctypes.c_int, [_h, ctypes.c_char, ctypes.c_double, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                                ctypes.c_int, ctypes.c_int]),
    "dmhm_hmat_multiply_dense_noalloc": (ctypes.c_int, [_h, ctypes.c_double, ctypes.c_void_p, ctypes.c_int,
                                                        ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                                        ctypes.c_int]),
    "dmhm_hmat_multiply_hmat": (ctypes.c_int, [ctypes.c_double, _h, _h, _hp, ctypes.POINTER(HHOpts)]),
    "dmhm_hmat_multiply_hmat_update": (ctypes.c_int, [ctypes.c_double, _h, _h, ctypes.c_double, _h,
                                                      ctypes.POINTER(HHOpts)]),
    "dmhm_hmat_update_with": (ctypes.c_int, [_h, ctypes.c_double, _h, ctypes.POINTER(HHOpts)]),
    "dmhm_hmat_create_from_lowrank": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int, _dp, _dp, _hp]),
    "dmhm_rounded_update": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]),
    "dmhm_compress_dense": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip]),
    "dmhm_profile_hh_phases": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _ip, _ip]),
    "dmhm_set_tile_engine": (ctypes.c_int, [ctypes.c_int]),
    "dmhm_rank_model_create": (ctypes.c_int, [ctypes.c_uint64, _hp]),
    "dmhm_rank_model_destroy": (ctypes.c_int, [_h]),
    "dmhm_rank_model_num_params": (ctypes.c_int, []),
    "dmhm_rank_model_get_params": (ctypes.c_int, [_h, _fp, ctypes.c_int]),
    "dmhm_rank_model_set_params": (ctypes.c_int, [_h, _fp, ctypes.c_int]),
    "dmhm_rank_model_gradient": (ctypes.c_int, [_h, _fp, _fp, ctypes.c_int, _fp, _fp]),
    "dmhm_rank_model_predict": (ctypes.c_int, [_h, _fp, ctypes.c_int, _fp]),
    "dmhm_rank_model_train": (ctypes.c_int, [_h, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_float, ctypes.c_uint64, _fp]),
    "dmhm_hmat_rank_features": (ctypes.c_int, [_h, _fp, ctypes.c_int]),
    "dmhm_last_product_stats": (ctypes.c_int, [_llp]),
    "dmhm_get_tile_engine": (ctypes.c_int, []),
    "dmhm_hmat_packed_size": (ctypes.c_int, [_h, ctypes.POINTER(ctypes.c_uint64)]),
    "dmhm_hmat_pack": (ctypes.c_int, [_h, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "dmhm_hmat_unpack": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _hp]),
    "dmhm_hmat_unpack_image": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), _hp]),
    "dmhm_hmat_copy": (ctypes.c_int, [_h, _hp]),
    "dmhm_hmat_transpose": (ctypes.c_int, [_h, _hp]),
    "dmhm_hmat_scale": (ctypes.c_int, [_h, ctypes.c_double]),
    "dmhm_hmat_set_to_identity": (ctypes.c_int, [_h]),
    "dmhm_hmat_add_constant_to_diagonal": (ctypes.c_int, [_h, ctypes.c_double]),
    "dmhm_hmat_convert": (ctypes.c_int, [_h, ctypes.c_int, _hp]),
    "dmhm_hmat_estimate_two_norm": (ctypes.c_int, [_h, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _dp]),
    "dmhm_hmat_schulz_invert": (ctypes.c_int, [_h, ctypes.POINTER(SchulzOpts), ctypes.POINTER(SchulzReport)]),
    "dmhm_hmat_direct_invert": (ctypes.c_int, [_h, ctypes.POINTER(HHOpts)]),
    "dmhm_hbatch_create": (ctypes.c_int, [ctypes.c_int, _hp, _hp]),
    "dmhm_hbatch_destroy": (ctypes.c_int, [_h]),
    "dmhm_hbatch_count": (ctypes.c_int, [_h, _ip]),
    "dmhm_hbatch_member": (ctypes.c_int, [_h, ctypes.c_int, _hp]),
    "dmhm_hbatch_multiply_hmat": (ctypes.c_int, [ctypes.c_double, _h, _h, ctypes.POINTER(HHOpts), _hp]),
    "dmhm_hbatch_estimate_two_norm": (ctypes.c_int, [_h, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _dp]),
    "dmhm_hbatch_schulz_invert": (ctypes.c_int, [_h, ctypes.POINTER(SchulzOpts), ctypes.POINTER(SchulzReport)]),
    "dmhm_direct_invert_program_stats": (ctypes.c_int, [_h, ctypes.c_int, _llp, _llp, _dp]),
    "dmhm_check_direct_invert_program": (ctypes.c_int, [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                         ctypes.c_int, _llp]),
    "dmhm_hmat_to_dense": (ctypes.c_int, [_h, _dp, ctypes.c_int]),
    "dmhm_hmat_ranks": (ctypes.c_int, [_h, _ip, ctypes.c_int]),
    "dmhm_hmat_matvec_bytes": (ctypes.c_int, [_h, ctypes.c_int, _dp, _dp]),
    "dmhm_profile_matvec": (ctypes.c_int, [_h, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "dmhm_profile_hh": (ctypes.c_int, [ctypes.c_double, _h, _h, ctypes.POINTER(HHOpts), _dp, _dp]),
}

_lib = None


def load():
    """Load libdmhm_amd.so and bind every exported entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libdmhm_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code == DMHM_OK:
        return
    msg = load().dmhm_last_error().decode(errors="replace")
    if code == 2:
        raise NonConformalError(code, msg)
    if code == 4:
        raise NotConvergedError(code, msg)
    raise DmhmError(code, msg)


_initialized = False


TILE_ENGINES = {"auto": -1, "lds": 0, "mfma": 1, "mfma_pf": 2, "bf16": 3, "bf16_pf": 4}


def set_tile_engine(name):
    """Force a tile engine ("lds", "mfma", "mfma_pf", "bf16", "bf16_pf") or
    autotune ("auto", the default); returns the previous setting's name."""
    prev = load().dmhm_get_tile_engine()
    check(load().dmhm_set_tile_engine(TILE_ENGINES[name]))
    return {v: k for k, v in TILE_ENGINES.items()}[prev]


def init(device=0):
    """dmhm_init once per process (binds the GPU)."""
    global _initialized
    if not _initialized:
        check(load().dmhm_init(device))
        _initialized = True
