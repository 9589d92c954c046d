"""bgk_step_diagnostics: time steps followed by the diagnostics of the state
they leave, the last two steps and the reductions fused into one pass.

* Populations bit-identical to bgk_step's (checksums and full arrays), every
  storage kind, ragged shapes, partial tiles and z chunks.
* Sums equal to bgk_step + bgk_diagnostics to FP64 rounding (the same
  per-site values summed in another order): means within 1e-14 absolute,
  dissipation sums within 1e-12 relative.
* Against the CPU oracle (diagno and dissipation of the oracle's state after
  the same steps): 1e-12 on diagno's means, 1e-10 relative on dissipation,
  the bar of the separate reductions.
* The fused pass is the one that runs where it should (timing counters), and
  the separate path where it should not (3k steps, halo images).
"""
import sys

import numpy as np
import pytest

import bgk3d_hit_amd as bgk

pytestmark = pytest.mark.gpu

OMEGA = 1.93


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def pair(shape, kind, seed=11, refcompat=False, two_step="always"):
    u = bgk.hit_velocity(*shape, 3.0, 0.04, seed)
    lats = []
    for _ in range(2):
        lat = bgk.Lattice(*shape, omega=OMEGA, storage=kind, timing=True,
                          refcompat_halos=refcompat, two_step=two_step)
        lat.init_equilibrium(u)
        lats.append(lat)
    return lats


def check_sums(d_f, s_f, d_s, s_s):
    assert np.all(np.abs(d_f - d_s) <= 1e-14 + 1e-13 * np.abs(d_s)), (d_f, d_s)
    assert np.all(np.abs(s_f - s_s) <= 1e-12 * np.abs(s_s)), (s_f, s_s)


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix", "f16", "bf16"])
@pytest.mark.parametrize("shape", [(70, 36, 44), (64, 40, 67), (40, 36, 4), (33, 17, 5)])
def test_fused_equals_step_then_diagnostics(gpu, kind, shape):
    nu = bgk.svisc_from_omega(OMEGA)
    a, b = pair(shape, kind)
    try:
        for n in (2, 5, 4, 6, 8):
            say("step_diagnostics", kind, shape, n)
            d_f, s_f = a.step_diagnostics(n, nu)
            b.step(n)
            d_s, s_s = b.diagnostics(nu)
            assert a.checksum() == b.checksum()
            check_sums(d_f, s_f, d_s, s_s)
            fused = a.timing_fused(reset=True)[1]
            assert fused == (0 if n % 3 == 0 else 1), (n, fused)
        assert np.array_equal(a.get_populations(), b.get_populations())
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("kind", ["f64", "f32", "f16"])
def test_fused_against_oracle(gpu, kind):
    import oracle

    shape = (40, 36, 34)
    nu = bgk.svisc_from_omega(OMEGA)
    say("step_diagnostics vs oracle", kind, shape)
    u = bgk.hit_velocity(*shape, 3.0, 0.04, 5)
    g = bgk.Lattice(*shape, omega=OMEGA, storage=kind, timing=True, two_step="always")
    o = oracle.OracleLattice(*shape, omega=OMEGA, kind=kind)
    try:
        g.init_equilibrium(u)
        o.init_equilibrium(u)
        for n in (5, 7, 2):
            d5, s2 = g.step_diagnostics(n, nu)
            o.step(n)
            assert np.array_equal(g.get_populations(), o.get_populations())
            assert np.abs(d5 - o.diagno()).max() <= 1e-12
            o.bcond_step3()
            so = o.dissipation(nu)
            assert np.all(np.abs(s2 - so) <= 1e-10 * np.abs(so))
        assert g.timing_fused()[1] == 3
    finally:
        g.close()


def test_halo_images_take_the_separate_path(gpu):
    shape = (48, 40, 36)
    nu = bgk.svisc_from_omega(OMEGA)
    say("step_diagnostics with halo images", shape)
    a, b = pair(shape, "f32", refcompat=True)
    try:
        d_f, s_f = a.step_diagnostics(5, nu)
        b.step(5)
        d_s, s_s = b.diagnostics(nu)
        assert a.checksum() == b.checksum()
        assert np.array_equal(d_f, d_s) and np.array_equal(s_f, s_s)
        assert a.timing_fused()[1] == 0
    finally:
        a.close()
        b.close()


def test_one_step_takes_the_separate_path(gpu):
    shape = (48, 40, 36)
    nu = bgk.svisc_from_omega(OMEGA)
    a, b = pair(shape, "f32")
    try:
        d_f, s_f = a.step_diagnostics(1, nu)
        b.step(1)
        d_s, s_s = b.diagnostics(nu)
        assert a.checksum() == b.checksum()
        assert np.array_equal(d_f, d_s) and np.array_equal(s_f, s_s)
        assert a.timing_fused()[1] == 0
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("kind", ["f32", "f16"])
def test_512cubed_window(gpu, kind):
    """The timed window of the headline workload: 20 steps and the
    diagnostics, at full size and the default pass selection."""
    shape = (512, 512, 512)
    nu = bgk.svisc_from_omega(1.95)
    say("512^3 step_diagnostics(20)", kind)
    a = bgk.Lattice(*shape, omega=1.95, storage=kind, timing=True)
    try:
        a.init_hit(4.0, 0.02, 1234)
        d_f, s_f = a.step_diagnostics(20, nu)
        h_f = a.checksum()
        assert a.timing_passes(reset=False)[3][1] == 6
        assert a.timing_fused()[1] == 1
    finally:
        a.close()
    b = bgk.Lattice(*shape, omega=1.95, storage=kind)
    try:
        b.init_hit(4.0, 0.02, 1234)
        b.step(20)
        d_s, s_s = b.diagnostics(nu)
        assert b.checksum() == h_f
    finally:
        b.close()
    check_sums(d_f, s_f, d_s, s_s)


def test_argument_errors(gpu):
    """Negative step counts and a null output are refused before any work,
    and the state is left as it was."""
    import ctypes

    from bgk3d_hit_amd import _lib as L

    shape = (48, 40, 36)
    a, _ = pair(shape, "f32")
    _.close()
    try:
        h0 = a.checksum()
        out = (ctypes.c_double * 7)()
        assert a._lib.bgk_step_diagnostics(a._h, -1, 0.01, out) == L.EINVAL
        assert "nsteps" in L.last_error()
        assert a._lib.bgk_step_diagnostics(a._h, 5, 0.01, None) == L.EINVAL
        assert a.checksum() == h0
        with pytest.raises(L.BGKError):
            a.step_diagnostics(-3, 0.01)
        # zero steps: the diagnostics of the current state
        d0, s0 = a.step_diagnostics(0, 0.01)
        d1, s1 = a.diagnostics(0.01)
        assert np.array_equal(d0, d1) and np.array_equal(s0, s1)
        assert a.checksum() == h0
    finally:
        a.close()


def test_512cubed_2000_steps_diagnostics_every_50(gpu):
    """The full headline run (512^3 FP32, omega 1.95, 2000 steps, the
    diagnostics every 50 steps): 40 fused calls against 40 x (bgk_step(50) +
    bgk_diagnostics), every diagnostic compared, the final states equal."""
    shape = (512, 512, 512)
    nu = bgk.svisc_from_omega(1.95)
    say("512^3 FP32 2000 steps, diagnostics every 50")
    a = bgk.Lattice(*shape, omega=1.95, storage="f32", timing=True)
    b = bgk.Lattice(*shape, omega=1.95, storage="f32")
    try:
        a.init_hit(4.0, 0.02, 1234)
        b.init_hit(4.0, 0.02, 1234)
        ke0 = a.diagno()[4]
        for blk in range(40):
            d_f, s_f = a.step_diagnostics(50, nu)
            b.step(50)
            d_s, s_s = b.diagnostics(nu)
            check_sums(d_f, s_f, d_s, s_s)
        assert a.checksum() == b.checksum()
        assert a.timing_fused()[1] == 40
        # decaying turbulence: energy falls, mass stays
        assert d_f[4] < ke0
        assert abs(d_f[0] - 1.0) < 1e-6
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("kind", ["f64", "f32", "f16"])
def test_unshifted_storage(gpu, kind):
    """Populations stored as f_i (BGK_FLAG_NOSHIFT): the fused pass's
    unshifted instantiation against bgk_step + bgk_diagnostics and the oracle."""
    import oracle

    shape = (44, 30, 26)
    nu = bgk.svisc_from_omega(OMEGA)
    say("step_diagnostics unshifted", kind, shape)
    u = bgk.hit_velocity(*shape, 3.0, 0.04, 9)
    a = bgk.Lattice(*shape, omega=OMEGA, storage=kind, shifted=False, timing=True,
                    two_step="always")
    b = bgk.Lattice(*shape, omega=OMEGA, storage=kind, shifted=False, two_step="always")
    o = oracle.OracleLattice(*shape, omega=OMEGA, kind=kind, shifted=False)
    try:
        for lat in (a, b, o):
            lat.init_equilibrium(u)
        d_f, s_f = a.step_diagnostics(8, nu)
        b.step(8)
        d_s, s_s = b.diagnostics(nu)
        o.step(8)
        assert a.checksum() == b.checksum()
        assert np.array_equal(a.get_populations(), o.get_populations())
        check_sums(d_f, s_f, d_s, s_s)
        assert np.abs(d_f - o.diagno()).max() <= 1e-12
        o.bcond_step3()
        so = o.dissipation(nu)
        assert np.all(np.abs(s_f - so) <= 1e-10 * np.abs(so))
        assert a.timing_fused()[1] == 1
    finally:
        a.close()
        b.close()
