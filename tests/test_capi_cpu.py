"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, validates arguments before touching a GPU, and its host
HIT generator meets its specification.  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

import bgk3d_hit_amd as bgk
from bgk3d_hit_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "bgk3d_hit.h")).read()
    return sorted(set(re.findall(r"^int\s+(bgk_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "ctypes table and header disagree"


def test_version_and_device_count():
    lib = L.load()
    assert lib.bgk_version() >= 100
    n = ctypes.c_int(-1)
    assert lib.bgk_device_count(ctypes.byref(n)) == L.OK
    assert n.value >= 0


@pytest.mark.parametrize("args,code", [
    ((0, 8, 8, 0, 0, 0, 1.9), L.EINVAL),        # empty lattice
    ((8, 8, 8, 9, 0, 0, 1.9), L.EINVAL),        # unknown storage kind
    ((8, 8, 8, 0, 5, 0, 1.9), L.EINVAL),        # unknown layout
    ((8, 8, 8, 0, 0, 0, 2.0), L.EINVAL),        # omega outside [0, 2)
    ((8, 8, 8, 0, 0, 0, -0.1), L.EINVAL),
    ((8, 8, 1, 0, 1, 0, 1.9), L.EINVAL),        # slab needs 2 planes
    ((8, 8, 8, 0, 1, 2, 1.9), L.EINVAL),        # refcompat images: periodic only
    ((2048, 2048, 1024, 0, 0, 0, 1.9), L.EINVAL),  # > 2^31 sites per population
])
def test_create_rejects_invalid_arguments(args, code):
    lib = L.load()
    h = ctypes.c_void_p()
    assert lib.bgk_create(*args, ctypes.byref(h)) == code
    assert not h.value
    assert len(L.last_error()) > 0


@pytest.mark.parametrize("shape,kind,flags", [
    ((4224, 4096, 64), L.STORE_F32, 0),   # 33 * 1024 FP32 reduction tiles per plane
    ((4096, 4096, 8), L.STORE_F64, 0),    # 64 * 1024 FP64 reduction tiles per plane
    ((0, 8, 8), L.STORE_F32, 0),          # empty lattice
    ((8, 8, 8), 9, 0),                    # unknown storage kind
    ((8, 8, 8), L.STORE_F32, L.FLAG_REFCOMPAT_HALOS),  # NOSHIFT is the only flag
    ((2048, 2048, 1024), L.STORE_F16, 0),  # > 2^31 sites per population
])
def test_noalloc_diagnostics_rejects_before_touching_a_gpu(shape, kind, flags):
    lib = L.load()
    pop, out = ctypes.c_void_p(256), ctypes.c_void_p(4096)  # never dereferenced
    assert lib.bgk_diagnostics_noalloc(pop, *shape, kind, flags, 0.01, out, None) == L.EINVAL
    assert len(L.last_error()) > 0
    assert lib.bgk_diagnostics_noalloc(None, 8, 8, 8, L.STORE_F32, 0, 0.01, out, None) == \
        L.EINVAL


def test_create_without_gpu_reports_no_device():
    if bgk.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(bgk.BGKError) as e:
        bgk.Lattice(8, 8, 8, omega=1.9)
    assert e.value.code == L.ENODEV


def test_null_context_is_an_error():
    lib = L.load()
    assert lib.bgk_col_MC(None) == L.EINVAL
    assert lib.bgk_destroy(None) == L.OK


def test_svisc_omega_round_trip():
    """SRC/hencol.F90:31 and its inverse."""
    for om in (1.0, 1.9, 1.95, 1.97):
        assert abs(bgk.omega_from_svisc(bgk.svisc_from_omega(om)) - om) < 1e-15
    assert abs(bgk.omega_from_svisc(0.0014317) - 2 / (6 * 0.0014317 + 1)) < 1e-16


# ------------------------------------------------------------ HIT generator
def spectrum(u):
    lz, ly, lx = u.shape[1:]
    return np.fft.fftn(u, axes=(1, 2, 3)) / (lx * ly * lz)


def test_hit_basic_statistics():
    n, k0, urms = 32, 3.0, 0.02
    u = bgk.hit_velocity(n, n, n, k0, urms, 1234)
    assert u.shape == (3, n, n, n)
    assert np.abs(u.mean(axis=(1, 2, 3))).max() < 1e-17
    assert abs(np.sqrt((u * u).sum(axis=0).mean() / 3) - urms) < 1e-15


def test_hit_is_solenoidal_and_follows_spectrum():
    n, k0, urms = 32, 3.0, 0.02
    u = bgk.hit_velocity(n, n, n, k0, urms, 99)
    U = spectrum(u)
    k = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    div = kx * U[0] + ky * U[1] + kz * U[2]
    assert np.abs(div).max() < 1e-14 * np.abs(U).max()
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    amp2 = (np.abs(U) ** 2).sum(axis=0)
    inband = (kk > 0) & (kk <= 4 * k0) & (np.abs(kx) < n / 2) & (np.abs(ky) < n / 2) & \
             (np.abs(kz) < n / 2)
    assert amp2[~inband].max() < 1e-30
    # |u^(k)|^2 = s^2 |kappa|^2 exp(-2 |kappa|^2/k0^2) for every mode: shell energy
    # sum over ~4 pi k^2 modes then follows k^4 exp(-2 (k/k0)^2)
    model = kk[inband] ** 2 * np.exp(-2 * (kk[inband] / k0) ** 2)
    ratio = amp2[inband] / model
    assert ratio.std() < 1e-10 * ratio.mean()
    # real field: Hermitian symmetry
    assert np.abs(u.imag if np.iscomplexobj(u) else 0).max() == 0


def test_hit_deterministic_and_seeded():
    a = bgk.hit_velocity(16, 16, 16, 2.0, 0.03, 5)
    b = bgk.hit_velocity(16, 16, 16, 2.0, 0.03, 5)
    c = bgk.hit_velocity(16, 16, 16, 2.0, 0.03, 6)
    assert np.array_equal(a, b)
    assert np.abs(a - c).max() > 1e-3


def test_hit_same_field_at_two_resolutions():
    """Per-mode counter-based draws: when both lattices resolve every mode of
    the band, the Fourier coefficients agree."""
    a = bgk.hit_velocity(24, 24, 24, 1.5, 0.02, 11)
    b = bgk.hit_velocity(48, 48, 48, 1.5, 0.02, 11)
    A, B = spectrum(a), spectrum(b)
    for kx, ky, kz in [(1, 0, 0), (2, -3, 1), (0, 4, -2), (5, 1, 1)]:
        assert np.allclose(A[:, kz, ky, kx], B[:, kz, ky, kx], rtol=1e-12, atol=1e-18)
    # and b restricted to the coarse grid is a (the field is band limited)
    assert np.abs(b[:, ::2, ::2, ::2] - a).max() < 1e-15


def test_hit_anisotropic_box():
    u = bgk.hit_velocity(32, 32, 8, 2.0, 0.03, 5678)
    assert abs(np.sqrt((u * u).sum(axis=0).mean() / 3) - 0.03) < 1e-15
    U = spectrum(u)
    kx = np.fft.fftfreq(32, 1 / 32)[None, None, :]
    ky = np.fft.fftfreq(32, 1 / 32)[None, :, None]
    kz = np.fft.fftfreq(8, 1 / 8)[:, None, None] * 4  # kappa_z = kz * lx / lz
    div = kx * U[0] + ky * U[1] + kz * U[2]
    assert np.abs(div).max() < 1e-14 * np.abs(U).max()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout + r.stderr
