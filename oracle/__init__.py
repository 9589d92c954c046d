"""CPU oracle of the BGK3D_HIT hot path -- TEST INFRASTRUCTURE ONLY.

A plain-C restatement of the reference Fortran (liboracle.so, built from
oracle/bgk_oracle.c by oracle/Makefile) with a thin ctypes wrapper.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as
the checker; the product library never loads it.

Reference build status: the reference's own sources cannot be compiled here
without writing a stand-in `mpi` module (SRC/precision.F90:3 and
SRC/storage.F90:28 `use mpi`; the only mpi.mod in the image is in gfortran
format, which amdflang cannot read), so there is no oracle/_ref build.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

F64, F32, F32MIX, F16, BF16 = range(5)
KINDS = {"f64": F64, "f32": F32, "f32mix": F32MIX, "f16": F16, "bf16": BF16}

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def build():
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "orc_create": (vp, [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_set_omega": (None, [vp, ctypes.c_double]),
        "orc_set_populations": (None, [vp, _dp]),
        "orc_get_populations": (None, [vp, _dp]),
        "orc_get_halo_array": (None, [vp, ctypes.c_int, _dp]),
        "orc_init_equilibrium": (None, [vp, _dp]),
        "orc_bcond_step3": (None, [vp]),
        "orc_col_MC": (None, [vp]),
        "orc_col_MC_masked": (None, [vp, ctypes.c_int, ctypes.c_int]),
        "orc_step": (None, [vp, ctypes.c_int]),
        "orc_diagno": (None, [vp, ctypes.c_int, ctypes.c_int, _dp]),
        "orc_dissipation": (None, [vp, ctypes.c_double, _dp]),
        "orc_adjoint_ke": (ctypes.c_double, [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int,
                                                                  _dp, _dp]),
        "orc_forward_ke": (ctypes.c_double, [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int,
                                                                  _dp]),
        "orc_f32_to_f16": (ctypes.c_uint16, [ctypes.c_float]),
        "orc_f16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
        "orc_f32_to_bf16": (ctypes.c_uint16, [ctypes.c_float]),
        "orc_bf16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
        "orc_num_threads": (ctypes.c_int, []),
        "orc_set_num_threads": (None, [ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


class OracleLattice:
    """Reference-layout lattice (halo'd arrays, pointer swap) on the CPU."""

    def __init__(self, lx, ly, lz, omega, kind="f64", shifted=True):
        self._lib = load()
        k = KINDS[kind] if isinstance(kind, str) else int(kind)
        self._h = self._lib.orc_create(lx, ly, lz, k, float(omega), 1 if shifted else 0)
        if not self._h:
            raise MemoryError("orc_create failed")
        self.shape = (lx, ly, lz)

    def close(self):
        if self._h:
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, n):
        lx, ly, lz = self.shape
        return np.empty((n, lz, ly, lx))

    def set_omega(self, omega):
        self._lib.orc_set_omega(self._h, float(omega))

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        self._lib.orc_set_populations(self._h, _d(f))

    def get_populations(self):
        f = self._arr(19)
        self._lib.orc_get_populations(self._h, _d(f))
        return f

    def halo_array(self, q):
        lx, ly, lz = self.shape
        a = np.empty((lz + 2, ly + 2, lx + 2))
        self._lib.orc_get_halo_array(self._h, int(q), _d(a))
        return a

    def init_equilibrium(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._lib.orc_init_equilibrium(self._h, _d(u))

    def bcond_step3(self):
        self._lib.orc_bcond_step3(self._h)

    def col_MC(self):
        self._lib.orc_col_MC(self._h)

    def col_MC_masked(self, opt, border):
        self._lib.orc_col_MC_masked(self._h, int(opt), int(border))

    def step(self, n=1):
        self._lib.orc_step(self._h, int(n))

    def diagno(self, ztot_bug=False, acc_mykind=False):
        out = np.zeros(5)
        self._lib.orc_diagno(self._h, int(ztot_bug), int(acc_mykind), _d(out))
        return out

    def dissipation(self, svisc):
        out = np.zeros(2)
        self._lib.orc_dissipation(self._h, float(svisc), _d(out))
        return out


def adjoint_ke(lx, ly, lz, omega, nsteps, u0):
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    g = np.empty_like(u0)
    J = load().orc_adjoint_ke(lx, ly, lz, float(omega), int(nsteps), _d(u0), _d(g))
    return J, g


def forward_ke(lx, ly, lz, omega, nsteps, u0):
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    return load().orc_forward_ke(lx, ly, lz, float(omega), int(nsteps), _d(u0))


def num_threads():
    return load().orc_num_threads()


def set_num_threads(n):
    load().orc_set_num_threads(int(n))
