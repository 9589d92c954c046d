"""Python host mirror of the reference's hot-path subroutines.

``Lattice`` owns one device context of the C-ABI.  Its methods carry the
names of the reference subroutines they replace so that a time loop reads
like SRC/bgk3d_HIT.F90:127-146::

    lat = Lattice(64, 64, 64, omega=1.9, storage="f64")
    lat.init_hit(k0=4.0, urms=0.02, seed=1234)
    for itime in range(1, itfin + 1):
        lat.boundaries()          # SRC/boundaries.F90:23
        lat.col_MC()              # SRC/col_MC.F90:34
        if itime % icheck == 0:
            lat.diagno()          # SRC/diagno.F90:28
            lat.dissipation(svisc)  # SRC/dissipation.F90:22
"""
import ctypes

import numpy as np

from . import _lib as L


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def omega_from_svisc(svisc):
    """SRC/hencol.F90:31: omega = 2 / (6 nu + 1)."""
    return 2.0 / (6.0 * svisc + 1.0)


def svisc_from_omega(omega):
    """Inverse of SRC/hencol.F90:31: nu = (2/omega - 1) / 6."""
    return (2.0 / omega - 1.0) / 6.0


def hit_velocity(lx, ly, lz, k0, urms, seed):
    """Host HIT velocity field, shape (3, lz, ly, lx) (x fastest in memory)."""
    u = np.empty((3, lz, ly, lx), dtype=np.float64)
    L.check(L.load().bgk_hit_velocity(lx, ly, lz, float(k0), float(urms), int(seed), _dptr(u)))
    return u


def device_count():
    n = ctypes.c_int(0)
    L.check(L.load().bgk_device_count(ctypes.byref(n)))
    return n.value


class Lattice:
    """One device lattice (SoA populations resident in HBM).

    two_step: True fuses two time steps per pass in step() where the lattice
    is large enough for it to pay, False takes one pass per step, "always"
    fuses at any size (testing).  three_step: with two_step, also fuse three
    time steps per pass (periodic layout) where two_step would fuse two."""

    def __init__(self, lx, ly, lz, omega, storage="f64", layout="periodic", shifted=True,
                 refcompat_halos=False, timing=False, two_step=True, wide_offsets=False,
                 three_step=True):
        self._lib = L.load()
        kind = L.STORAGE_NAMES[storage] if isinstance(storage, str) else int(storage)
        lay = {"periodic": L.LAYOUT_PERIODIC, "slab": L.LAYOUT_SLAB_Z}[layout]
        flags = (0 if shifted else L.FLAG_NOSHIFT) | \
                (L.FLAG_REFCOMPAT_HALOS if refcompat_halos else 0) | \
                (L.FLAG_TIMING if timing else 0) | \
                {True: 0, False: L.FLAG_SINGLE_STEP, "always": L.FLAG_TWO_STEP_ANY_SIZE}[two_step] | \
                (L.FLAG_WIDE_OFFSETS if wide_offsets else 0) | \
                (0 if three_step else L.FLAG_NO_THREE_STEP)
        h = ctypes.c_void_p()
        L.check(self._lib.bgk_create(lx, ly, lz, kind, lay, flags, float(omega), ctypes.byref(h)))
        self._h = h
        self.shape = (lx, ly, lz)
        self.omega = float(omega)
        self.storage = kind

    # -- lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.bgk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def nsite(self):
        lx, ly, lz = self.shape
        return lx * ly * lz

    def _arr(self, ncomp):
        lx, ly, lz = self.shape
        return np.empty((ncomp, lz, ly, lx), dtype=np.float64)

    def _in(self, a, ncomp):
        lx, ly, lz = self.shape
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != ncomp * lx * ly * lz:
            raise ValueError(f"expected {ncomp} x {lz} x {ly} x {lx} values, got {a.shape}")
        return a

    # -- configuration
    def set_omega(self, omega):
        L.check(self._lib.bgk_set_omega(self._h, float(omega)))
        self.omega = float(omega)

    def set_border(self, border):
        L.check(self._lib.bgk_set_border(self._h, int(border)))

    def info(self):
        v = (ctypes.c_int64 * 8)()
        L.check(self._lib.bgk_info(self._h, v))
        keys = ("lx", "ly", "lz", "storage", "layout", "flags", "device_bytes", "steps")
        return dict(zip(keys, list(v)))

    # -- state transfer
    def set_populations(self, f):
        f = self._in(f, 19)
        L.check(self._lib.bgk_set_populations(self._h, _dptr(f)))

    def get_populations(self):
        f = self._arr(19)
        L.check(self._lib.bgk_get_populations(self._h, _dptr(f)))
        return f

    def init_equilibrium(self, u):
        """SRC/init.F90:97-137 with rho = 1 from velocity u (3, lz, ly, lx)."""
        u = self._in(u, 3)
        L.check(self._lib.bgk_init_equilibrium(self._h, _dptr(u)))

    def init_hit(self, k0, urms, seed):
        L.check(self._lib.bgk_init_hit(self._h, float(k0), float(urms), int(seed)))

    # -- hot path (reference subroutine names)
    def boundaries(self):
        L.check(self._lib.bgk_boundaries(self._h))

    def col_MC(self):
        L.check(self._lib.bgk_col_MC(self._h))

    def col_MC_masked(self, opt):
        L.check(self._lib.bgk_col_MC_masked(self._h, int(opt)))

    def step(self, nsteps=1):
        L.check(self._lib.bgk_step(self._h, int(nsteps)))

    def diagno(self, refcompat=False):
        """(mean rho, mean ux, mean uy, mean uz, mean |u|^2/2)."""
        out = np.zeros(5)
        L.check(self._lib.bgk_diagno(self._h, L.DIAG_REFCOMPAT if refcompat else
                                     L.DIAG_CONSISTENT, _dptr(out)))
        return out

    def dissipation(self, svisc, refcompat=False):
        """(sum 2 svisc Pi:Pi, sum |u|^2/2) over pulled moments."""
        out = np.zeros(2)
        L.check(self._lib.bgk_dissipation(self._h, float(svisc), L.DIAG_REFCOMPAT if refcompat
                                          else L.DIAG_CONSISTENT, _dptr(out)))
        return out

    def diagnostics(self, svisc, refcompat=False):
        """diagno() and dissipation(svisc) in one pass over the populations
        (bit-identical to the two calls): returns (5 values, 2 values)."""
        out = np.zeros(7)
        L.check(self._lib.bgk_diagnostics(self._h, float(svisc), L.DIAG_REFCOMPAT if refcompat
                                          else L.DIAG_CONSISTENT, _dptr(out)))
        return out[:5].copy(), out[5:].copy()

    def step_diagnostics(self, nsteps, svisc):
        """step(nsteps), then diagnostics(svisc) of the state reached
        (consistent mode); the last two steps and the reductions share one
        pass where the lattice allows it.  Returns (5 values, 2 values)."""
        out = np.zeros(7)
        L.check(self._lib.bgk_step_diagnostics(self._h, int(nsteps), float(svisc), _dptr(out)))
        return out[:5].copy(), out[5:].copy()

    def get_box(self, x0, x1, y0, y1, z0, z1):
        """Populations of a sub-box, shape (19, z1-z0, y1-y0, x1-x0)."""
        out = np.empty((19, z1 - z0, y1 - y0, x1 - x0))
        L.check(self._lib.bgk_get_box(self._h, x0, x1, y0, y1, z0, z1, _dptr(out)))
        return out

    def checksum(self):
        """64-bit order-independent hash of the stored populations."""
        v = ctypes.c_uint64(0)
        L.check(self._lib.bgk_checksum(self._h, ctypes.byref(v)))
        return v.value

    def synchronize(self):
        L.check(self._lib.bgk_synchronize(self._h))

    def timing(self, reset=True):
        t = np.zeros(8)
        L.check(self._lib.bgk_timing(self._h, _dptr(t), 1 if reset else 0))
        return {"collide_ms": t[0], "collide_launches": int(t[1]), "reduce_ms": t[2],
                "reduce_launches": int(t[3]), "halo_ms": t[4], "halo_launches": int(t[5]),
                "step_ms": t[6], "step_calls": int(t[7])}

    def timing_passes(self, reset=True):
        """Collide launches split by the time steps one launch advances:
        {steps_per_launch: (ms, launches)} for 1, 2 and 3."""
        t = np.zeros(6)
        L.check(self._lib.bgk_timing_passes(self._h, _dptr(t), 1 if reset else 0))
        return {k + 1: (t[2 * k], int(t[2 * k + 1])) for k in range(3)}

    def timing_fused(self, reset=True):
        """Kernel time of the two-steps-plus-diagnostics passes of
        step_diagnostics (BGK_FLAG_TIMING): (ms, launches)."""
        t = np.zeros(2)
        L.check(self._lib.bgk_timing_fused(self._h, _dptr(t), 1 if reset else 0))
        return t[0], int(t[1])

    def timing_adjoint(self, reset=True):
        """Kernel time of the adjoint's taped forward passes and adjoint-step
        passes (BGK_FLAG_TIMING): {"tape": (ms, launches), "adjoint": (ms,
        launches)}."""
        t = np.zeros(4)
        L.check(self._lib.bgk_timing_adjoint(self._h, _dptr(t), 1 if reset else 0))
        return {"tape": (t[0], int(t[1])), "adjoint": (t[2], int(t[3]))}

    # -- adjoint
    def adjoint_ke(self, u0, nsteps, ckpt_every=25, levels=1):
        """(J, dJ/du0, stats) for J = sum |u|^2/2 after nsteps from equilibrium(u0)."""
        u0 = self._in(u0, 3)
        g = self._arr(3)
        J = ctypes.c_double(0.0)
        st = np.zeros(4)
        L.check(self._lib.bgk_adjoint_ke(self._h, _dptr(u0), int(nsteps), int(ckpt_every),
                                         int(levels), ctypes.byref(J), _dptr(g), _dptr(st)))
        return J.value, g, {"peak_bytes": st[0], "forward_ms": st[1], "backward_ms": st[2],
                            "recomputed_steps": int(st[3])}

    def adjoint_ke_device(self, u0_ptr, grad_ptr, nsteps, ckpt_every=25, levels=1):
        """Device-pointer variant: u0_ptr / grad_ptr address 3*nsite doubles in
        GPU memory (e.g. torch tensors' data_ptr()).  The call returns after
        the gradient is complete."""
        J = ctypes.c_double(0.0)
        st = np.zeros(4)
        L.check(self._lib.bgk_adjoint_ke_device(self._h, ctypes.c_void_p(int(u0_ptr)), int(nsteps),
                                                int(ckpt_every), int(levels), ctypes.byref(J),
                                                ctypes.c_void_p(int(grad_ptr)), _dptr(st)))
        return J.value, {"peak_bytes": st[0], "forward_ms": st[1], "backward_ms": st[2],
                         "recomputed_steps": int(st[3])}

    def forward_ke(self, u0, nsteps):
        u0 = self._in(u0, 3)
        J = ctypes.c_double(0.0)
        L.check(self._lib.bgk_forward_ke(self._h, _dptr(u0), int(nsteps), ctypes.byref(J)))
        return J.value
