"""ctypes binding of libmarqov_amd.so (the C ABI declared in include/marqov_amd.h).

The library is the product path; importing this module fails loudly when the
built library is missing, and there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmarqov_amd.so")

MQ_OK = 0
MQ_ISING, MQ_EA_ISING, MQ_POTTS, MQ_HEISENBERG, MQ_PHI4 = 1, 2, 3, 4, 5
MQ_PROPOSE_FRESH, MQ_PROPOSE_SMALL_ANGLE = 0, 1
MQ_EMBED_CARTESIAN, MQ_EMBED_RANDOM_DIR = 0, 1
MQ_NOBS = 4


class Tuning(ctypes.Structure):
    _fields_ = [("kernels", ctypes.c_int), ("block", ctypes.c_int), ("xcd_remap", ctypes.c_int),
                ("lds_block", ctypes.c_int), ("wolff_block", ctypes.c_int), ("wolff_ring", ctypes.c_int),
                ("wolff_vset", ctypes.c_int), ("graph", ctypes.c_int)]


class ModelParams(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int),
        ("J", ctypes.c_double),
        ("q", ctypes.c_int),
        ("lambda_", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("proposal", ctypes.c_int),
        ("delta", ctypes.c_double),
        ("embedding", ctypes.c_int),
        ("spin_bits", ctypes.c_int),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_D = ctypes.c_double

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "mq_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "mq_ctx_destroy": (None, [_P]),
    "mq_last_error": (ctypes.c_char_p, [_P]),
    "mq_sync": (_I, [_P]),
    "mq_timer_begin": (_I, [_P]),
    "mq_timer_end": (_I, [_P, ctypes.POINTER(_D)]),
    "mq_lattice_regular": (_I, [_P, _I, _I, ctypes.POINTER(_P)]),
    "mq_lattice_csr": (_I, [_P, _I64, _P, _P, ctypes.POINTER(_P)]),
    "mq_rgg_build": (_I, [_P, _I, _I64, _D, _U64, _P, _P, _P, _I64, _P]),
    "mq_lattice_colors": (_I, [_P, ctypes.POINTER(_I), _P]),
    "mq_lattice_nsites": (_I64, [_P]),
    "mq_lattice_color_of": (_I, [_P, _P]),
    "mq_lattice_destroy": (None, [_P]),
    "mq_state_create": (_I, [_P, ctypes.POINTER(ModelParams), _I, _P, ctypes.POINTER(_P)]),
    "mq_state_destroy": (None, [_P]),
    "mq_state_set_beta": (_I, [_P, _P]),
    "mq_state_get_beta": (_I, [_P, _P]),
    "mq_state_random_bonds": (_I, [_P, _P, _D, _U64]),
    "mq_state_set_bonds": (_I, [_P, _I, _P]),
    "mq_state_set_samples": (_I, [_P, _P]),
    "mq_state_init": (_I, [_P, _U64]),
    "mq_state_upload": (_I, [_P, _I, _P, _I64]),
    "mq_state_download": (_I, [_P, _I, _P, _I64]),
    "mq_metropolis": (_I, [_P, _I, _U64, _U64]),
    "mq_wolff": (_I, [_P, _I, _U64, _U64, _P]),
    "mq_emcs": (_I, [_P, _I, _I, _U64, _U64, _P]),
    "mq_emcs_run": (_I, [_P, _I, _I, _I, _U64, _U64, _P]),
    "mq_measure": (_I, [_P, _P]),
    "mq_measure_noalloc": (_I, [_P, _P]),
    "mq_measure_ft": (_I, [_P, _P]),
    "mq_energy": (_I, [_P, _P]),
    "mq_renormalize": (_I, [_P]),
    "mq_pt_swap": (_I, [_P, _I, _I, _U64, _U64, _P]),
    "mq_pt_swap_pairs": (_I, [_P, _I, _I, _P, _U64, _U64, _P]),
    "mq_pt_get_temps": (_I, [_P, _P]),
    "mq_state_set_tuning": (_I, [_P, ctypes.POINTER(Tuning)]),
    "mq_state_get_tuning": (_I, [_P, ctypes.POINTER(Tuning)]),
    "mq_run_graph": (_I, [_P, _I, _I, _U64, _U64, _U64, _I, _I]),
}

_lib = None


def load():
    """Load libmarqov_amd.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libmarqov_amd.so is not built (run __graft_entry__.build() or "
            "make -C marqov_amd/csrc); the HIP path has no fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
