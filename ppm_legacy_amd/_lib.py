"""ctypes binding of the in-tree libppmg.so (the HIP library).

There is no CPU fallback: if the library is missing, or no HIP device is
visible when a device entry point is called, the call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libppmg.so")

PPMG_OK = 0
PPMG_ERROR_FATAL = -1         # ppm_param.h:372-375
PPMG_ERROR_ERROR = -2
KERNEL_BSP2 = 1
KERNEL_MP4 = 2
PREC_F64 = 0
PREC_F32C = 1
PREC_F32 = 2
MAP_GHOST_GET = 5          # ppm_param.h:187-188
MAP_GHOST_PUT = 6
# element kinds of a ghost-mapped field (the _d, _s, _i, _l, _dc, _sc specifics)
KIND_D, KIND_S, KIND_I, KIND_L, KIND_DC, KIND_SC = 0, 2, 3, 4, 5, 6
MEM_DEVICE = 0
MEM_HOST = 1

ERR_ALLOC = 101
ERR_PART_RANGE = 402
ERR_PART_LOST = 403
ERR_PART_UNASS = 405
ERR_INDEX_CORR = 410
ERR_ARGUMENT = 503
ERR_NOINIT = 903
ERR_DEVICE = 9100
ERR_FIELD_BOUNDS = 9101


class PpmgMesh(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("nm", ctypes.c_int * 3), ("min_phys", ctypes.c_double * 3),
                ("max_phys", ctypes.c_double * 3), ("bcdef", ctypes.c_int * 6)]


class PpmgField(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("lda", ctypes.c_int), ("lo", ctypes.c_int * 3),
                ("ext", ctypes.c_int * 3)]


class PpmgError(RuntimeError):
    def __init__(self, info, code, msg):
        super().__init__(f"ppmg info={info} code={code}: {msg}")
        self.info = info
        self.code = code
        self.msg = msg


# every exported entry point declared in include/ppmg.h
EXPORTS = [
    "ppmg_plan_create", "ppmg_plan_destroy", "ppmg_p2m", "ppmg_p2m_noalloc", "ppmg_m2p", "ppmg_m2p_prebin", "ppmg_p2m_adjoint", "ppmg_util_rank",
    "ppmg_impose_part_bc", "ppmg_rmsh_comp_weights", "ppmg_rmsh_remesh", "ppmg_rmsh_create_part",
    "ppmg_plan_set_timing", "ppmg_plan_timings", "ppmg_plan_sort_view", "ppmg_last_error", "ppmg_set_debug", "ppmg_device_sync", "ppmg_init",
    "ppmg_finalize", "ppmg_set_precision_mode", "ppmg_mktopo_field", "ppmg_topo_destroy",
    "ppmg_interp_p2m", "ppmg_interp_m2p", "ppmg_rmsh_comp_weights_reg", "ppmg_rmsh_remesh_reg",
    "ppmg_util_rank_reg", "ppmg_topo_info", "ppmg_impose_part_bc_reg", "ppmg_rmsh_create_part_reg", "ppmg_gen_jittered_lattice", "ppmg_gen_normals", "ppmg_gen_divfree_velocity",
    "ppmg_map_field_ghost", "ppmg_map_field_ghost_reg", "ppmg_map_field_ghost_kind", "ppmg_map_field_ghost_kind_reg",
    "ppmg_util_sort", "ppmg_util_sort_reg",
]

_lib = None


def lib():
    """Load libppmg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m ppm_legacy_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def last_error():
    code = ctypes.c_int()
    buf = ctypes.create_string_buffer(512)
    info = lib().ppmg_last_error(ctypes.byref(code), buf, 512)
    return info, code.value, buf.value.decode(errors="replace")


def check(info):
    """Raise PpmgError for a non-zero info."""
    if info != 0:
        i, code, msg = last_error()
        raise PpmgError(info, code, msg)
    return info
