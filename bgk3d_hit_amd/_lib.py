"""ctypes binding of libbgk3d_hit.so (the C-ABI declared in include/bgk3d_hit.h).

The library is the only compute path: there is no Python or CPU fallback.
Loading fails with an error when the built library is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbgk3d_hit.so")

# status codes
OK, EINVAL, EHIP, ENOMEM, ESTATE, ENONFINITE, ENODEV = range(7)
# storage kinds
STORE_F64, STORE_F32, STORE_F32_MIXED, STORE_F16, STORE_BF16 = range(5)
STORAGE_NAMES = {"f64": STORE_F64, "f32": STORE_F32, "f32mix": STORE_F32_MIXED,
                 "f16": STORE_F16, "bf16": STORE_BF16}
# layouts
LAYOUT_PERIODIC, LAYOUT_SLAB_Z = 0, 1
# flags
FLAG_NOSHIFT, FLAG_REFCOMPAT_HALOS, FLAG_TIMING, FLAG_SINGLE_STEP = 1, 2, 4, 8
FLAG_TWO_STEP_ANY_SIZE = 16
FLAG_WIDE_OFFSETS = 32
FLAG_NO_THREE_STEP = 64
# diagnostic modes
DIAG_CONSISTENT, DIAG_REFCOMPAT = 0, 1
# reduction slots of bgk_diagnostics_noalloc's fixed scratch array
NOALLOC_MAX_PARTS = 32768

_c_int, _c_double, _p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every function returns an int status
SIGNATURES = {
    "bgk_version": [],
    "bgk_last_error": [ctypes.c_char_p, _c_int],
    "bgk_device_count": [ctypes.POINTER(_c_int)],
    "bgk_create": [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double,
                   ctypes.POINTER(_p)],
    "bgk_destroy": [_p],
    "bgk_set_omega": [_p, _c_double],
    "bgk_set_border": [_p, _c_int],
    "bgk_info": [_p, ctypes.POINTER(ctypes.c_int64)],
    "bgk_set_populations": [_p, _dp],
    "bgk_get_populations": [_p, _dp],
    "bgk_init_equilibrium": [_p, _dp],
    "bgk_hit_velocity": [_c_int, _c_int, _c_int, _c_double, _c_double, ctypes.c_uint64, _dp],
    "bgk_init_hit": [_p, _c_double, _c_double, ctypes.c_uint64],
    "bgk_boundaries": [_p],
    "bgk_col_MC": [_p],
    "bgk_col_MC_masked": [_p, _c_int],
    "bgk_step": [_p, _c_int],
    "bgk_diagno": [_p, _c_int, _dp],
    "bgk_dissipation": [_p, _c_double, _c_int, _dp],
    "bgk_diagnostics": [_p, _c_double, _c_int, _dp],
    "bgk_step_diagnostics": [_p, _c_int, _c_double, _dp],
    "bgk_diagnostics_noalloc": [_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double, _p, _p],
    "bgk_get_box": [_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _dp],
    "bgk_checksum": [_p, ctypes.POINTER(ctypes.c_uint64)],
    "bgk_synchronize": [_p],
    "bgk_timing": [_p, _dp, _c_int],
    "bgk_timing_passes": [_p, _dp, _c_int],
    "bgk_timing_adjoint": [_p, _dp, _c_int],
    "bgk_timing_fused": [_p, _dp, _c_int],
    "bgk_adjoint_ke": [_p, _dp, _c_int, _c_int, _c_int, _dp, _dp, _dp],
    "bgk_adjoint_ke_device": [_p, _p, _c_int, _c_int, _c_int, _dp, _p, _dp],
    "bgk_forward_ke": [_p, _dp, _c_int, _dp],
}


class BGKError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"[status {code}] {msg}")
        self.code = code


_lib = None


def load():
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C bgk3d_hit_amd` "
            "(or __graft_entry__.build()); there is no fallback path")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = _c_int
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().bgk_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status):
    if status != OK:
        raise BGKError(status, last_error())
    return status
