"""bgk_diagnostics_noalloc (the allocation-free diagnostics pass over
populations the caller owns) against bgk_diagnostics on the same state,
through the C-ABI: bit-identical sums on two calls in a row, the second on
a smaller lattice (the fixed scratch array keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

import bgk3d_hit_amd as bgk
from bgk3d_hit_amd import _lib as L

pytestmark = pytest.mark.gpu

KINDS = ["f64", "f32", "f32mix", "f16", "bf16"]


class DeviceBuffer:
    """Device memory of the test, from the HIP runtime the library runs on
    (the library's libamdhip64 is already loaded; opening the same soname
    returns it)."""
    H2D, D2H = 1, 2

    def __init__(self, nbytes):
        L.load()
        self.hip = hip = ctypes.CDLL("libamdhip64.so.7")
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, self.H2D) == 0

    def download(self, dtype, count):
        a = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, self.D2H) == 0
        return a

    def free(self):
        if self.ptr.value:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def device_copy(f, kind):
    """One population copy in the library's device layout ([z][y][q][x]: the
    19 population rows of each y line side by side) from host doubles
    f (19, lz, ly, lx), whose values the storage kind holds exactly."""
    inter = np.ascontiguousarray(f.transpose(1, 2, 0, 3))
    if kind == "f64":
        a = inter
        back = a
    elif kind in ("f32", "f32mix"):
        a = inter.astype(np.float32)
        back = a.astype(np.float64)
    elif kind == "f16":
        a = inter.astype(np.float16)
        back = a.astype(np.float64)
    else:  # bf16: the upper half of the float32 bits
        bits = inter.astype(np.float32).view(np.uint32)
        assert not (bits & 0xFFFF).any(), "value not representable in bf16"
        a = (bits >> 16).astype(np.uint16)
        back = (a.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    assert np.array_equal(back, inter), "host conversion is not exact"
    buf = DeviceBuffer(a.nbytes)
    buf.upload(a)
    return buf


def noalloc(pop, shape, kind, svisc, flags=0):
    out = DeviceBuffer(7 * 8)
    try:
        out.upload(np.full(7, np.nan))
        L.check(L.load().bgk_diagnostics_noalloc(pop.ptr, *shape, L.STORAGE_NAMES[kind], flags,
                                                 float(svisc), out.ptr, None))
        return out.download(np.float64, 7)  # hipMemcpy waits for the null stream
    finally:
        out.free()


def reference(shape, kind, omega, steps, seed, shifted=True):
    """bgk_diagnostics of a HIT lattice after `steps` steps, and its state."""
    nu = bgk.svisc_from_omega(omega)
    with bgk.Lattice(*shape, omega=omega, storage=kind, shifted=shifted) as lat:
        lat.init_hit(4.0, 0.02, seed)
        lat.step(steps)
        d5, s2 = lat.diagnostics(nu)
        f = lat.get_populations()
    return np.concatenate([d5, s2]), f, nu


@pytest.mark.parametrize("kind", KINDS)
def test_noalloc_diagnostics_equal_bgk_diagnostics(gpu, kind):
    # the larger lattice first, then a smaller one: the second call's
    # partials overwrite fewer slots of the shared scratch array
    for shape, steps, seed in (((72, 40, 36), 7, 11), ((24, 20, 16), 5, 12)):
        ref, f, nu = reference(shape, kind, 1.9, steps, seed)
        print(f"noalloc {kind} {shape}", flush=True)
        pop = device_copy(f, kind)
        got = noalloc(pop, shape, kind, nu)
        pop.free()
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (kind, shape, got, ref)


def test_noalloc_diagnostics_unshifted(gpu):
    for shape in ((40, 36, 32), (16, 12, 10)):
        ref, f, nu = reference(shape, "f64", 1.7, 4, 21, shifted=False)
        print(f"noalloc noshift {shape}", flush=True)
        pop = device_copy(f, "f64")
        got = noalloc(pop, shape, "f64", nu, flags=L.FLAG_NOSHIFT)
        pop.free()
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (shape, got, ref)


def test_noalloc_diagnostics_repeated_calls_agree(gpu):
    shape = (64, 48, 40)
    ref, f, nu = reference(shape, "f32", 1.95, 6, 5)
    pop = device_copy(f, "f32")
    for _ in range(3):
        got = noalloc(pop, shape, "f32", nu)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    pop.free()
