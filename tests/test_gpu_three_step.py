"""GPU parity of the three-step pass (k_collide3): bgk_step's fusion of three
time steps per pass over the populations must give the populations of
single steps bit for bit -- every intermediate value is rounded to the
storage type as a store would round it.  Sizes leave partial tiles in x and
y, partial last z chunks, lattices narrower than a tile (wrap within one
tile) and step counts that end with a two-step pass or a single step."""
import sys

import numpy as np
import pytest

import bgk3d_hit_amd as bgk

pytestmark = pytest.mark.gpu

KINDS3 = ["f64", "f32", "f32mix", "f16", "bf16"]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", KINDS3)
@pytest.mark.parametrize("shifted", [True, False])
@pytest.mark.parametrize("shape", [(70, 9, 6), (64, 16, 130), (24, 20, 16), (130, 17, 65),
                                   (67, 23, 4), (5, 3, 7), (129, 31, 9), (1, 1, 8), (3, 70, 4)])
def test_three_step_pass_equals_single_steps(gpu, kind, shifted, shape):
    say("three-step vs single", kind, shifted, shape)
    u = bgk.hit_velocity(*shape, 2.0, 0.03, 13)
    a = bgk.Lattice(*shape, omega=1.93, storage=kind, shifted=shifted, two_step="always",
                    timing=True)
    b = bgk.Lattice(*shape, omega=1.93, storage=kind, shifted=shifted, two_step=False)
    a.init_equilibrium(u)
    b.init_equilibrium(u)
    done = 0
    for n in (3, 1, 5, 6, 4):
        a.step(n)
        for _ in range(n):
            b.col_MC()
        done += n
        assert np.array_equal(a.get_populations(), b.get_populations()), (n, done)
    # 3 | 1 | 3+2 | 3+3 | 2+2 (a remainder of one after three-step passes
    # goes to two two-step passes): eight launches for 19 steps
    t = a.timing_passes()
    assert (t[1][1], t[2][1], t[3][1]) == (1, 3, 4)
    a.close()
    b.close()


@pytest.mark.parametrize("kind", KINDS3)
def test_full_rounds_and_a_tail_equal_single_steps(gpu, kind):
    """320x520x40 on 256 CUs: the three-step launch takes a round of whole
    tile columns and the rest in chunks (64x10 tiles: 256 columns whole and
    4 in 8-plane chunks, 276 workgroups; FP64 compute, 32x10: 512 whole and
    8 in chunks, 552) and the two-step launch 1248 workgroups (128x5, four
    rounds and a tail of 224) or, in FP64 compute, 2080 (64x5, eight rounds
    and 32), so the block order maps full rounds and spreads a tail over
    the XCDs in one launch.  Twelve steps of each pass depth equal twelve
    single steps."""
    say("full rounds and a tail 320x520x40", kind)
    shape = (320, 520, 40)
    u = bgk.hit_velocity(*shape, 4.0, 0.02, 31)
    ref = bgk.Lattice(*shape, omega=1.95, storage=kind, two_step=False)
    ref.init_equilibrium(u)
    for _ in range(12):
        ref.col_MC()
    want = ref.checksum()
    ref.close()
    for three in (True, False):
        a = bgk.Lattice(*shape, omega=1.95, storage=kind, three_step=three, timing=True)
        a.init_equilibrium(u)
        a.step(12)
        t = a.timing_passes()
        assert t[3 if three else 2][1] == (4 if three else 6), t
        assert a.checksum() == want, three
        a.close()


@pytest.mark.parametrize("kind", KINDS3)
def test_three_step_equals_two_step_passes(gpu, kind):
    """Three-step passes against two-step passes (BGK_FLAG_NO_THREE_STEP) on
    a lattice where both engage by themselves: 24 steps, checksums equal."""
    say("three-step vs two-step 192x160x128", kind)
    shape = (192, 160, 128)
    u = bgk.hit_velocity(*shape, 4.0, 0.02, 21)
    a = bgk.Lattice(*shape, omega=1.95, storage=kind, timing=True)
    b = bgk.Lattice(*shape, omega=1.95, storage=kind, three_step=False, timing=True)
    a.init_equilibrium(u)
    b.init_equilibrium(u)
    a.step(24)
    b.step(24)
    assert a.timing()["collide_launches"] == 8
    assert b.timing()["collide_launches"] == 12
    assert a.checksum() == b.checksum()


@pytest.mark.parametrize("kind", ["f64", "f32", "f16"])
def test_three_step_512cubed_checksum(gpu, kind):
    """The headline lattice: 512^3, 12 steps in four three-step passes ==
    12 single steps, bit for bit."""
    say("three-step 512^3", kind)
    a = bgk.Lattice(512, 512, 512, omega=1.95, storage=kind, timing=True)
    a.init_hit(4.0, 0.02, 1234)
    a.step(12)
    ca = a.checksum()
    assert a.timing()["collide_launches"] == 4
    a.close()
    b = bgk.Lattice(512, 512, 512, omega=1.95, storage=kind, two_step=False)
    b.init_hit(4.0, 0.02, 1234)
    b.step(12)
    assert b.checksum() == ca
    b.close()


def test_three_step_with_diagnostics_between(gpu):
    """Diagnostics between fused passes read the state the single steps
    leave (FP32, 256^3, diagnostics every 10 steps = 3+3+2+2)."""
    say("three-step diagnostics 256^3")
    nu = bgk.svisc_from_omega(1.95)
    a = bgk.Lattice(256, 256, 256, omega=1.95, storage="f32")
    b = bgk.Lattice(256, 256, 256, omega=1.95, storage="f32", two_step=False)
    a.init_hit(4.0, 0.02, 8)
    b.init_hit(4.0, 0.02, 8)
    for _ in range(3):
        a.step(10)
        b.step(10)
        da, sa = a.diagnostics(nu)
        db, sb = b.diagnostics(nu)
        assert np.array_equal(da, db) and np.array_equal(sa, sb)


@pytest.mark.parametrize("n,launches", [(64, 20), (128, 7), (256, 7)])
def test_three_step_engages_where_two_step_does(gpu, n, launches):
    """20 steps: 64^3 takes single steps (too few workgroups to fill the
    GPU); 128^3 and up take six three-step passes and one two-step pass."""
    g = bgk.Lattice(n, n, n, omega=1.9, storage="f32", timing=True)
    g.step(20)
    assert g.timing()["collide_launches"] == launches


@pytest.mark.parametrize("n,depths", [(1, (1, 0, 0)), (2, (0, 1, 0)), (4, (0, 2, 0)),
                                      (7, (0, 2, 1)), (25, (0, 2, 7)), (50, (0, 1, 16))])
def test_step_split_into_passes(gpu, n, depths):
    """How bgk_step splits n steps: three-step passes, then a two-step pass
    for a remainder of two, two for a remainder of one (cheaper than a
    three-step pass and a single step), single steps only below two."""
    g = bgk.Lattice(128, 128, 128, omega=1.9, storage="f32", timing=True)
    g.step(n)
    t = g.timing_passes()
    assert (t[1][1], t[2][1], t[3][1]) == depths


@pytest.mark.parametrize("kind", KINDS3)
@pytest.mark.parametrize("shape", [(70, 9, 12), (33, 20, 13), (64, 16, 40), (130, 17, 29)])
def test_slab_three_step_equals_single_steps(gpu, kind, shape):
    """One rank's slab: the bulk three-step pass (planes 3..nz-4) with the
    boundary planes stepped three times through two plane buffers, one
    exchange per step -- identical to periodic single steps, at the thinnest
    slab it takes (12 planes), odd plane counts and ragged x/y."""
    say("slab three-step", kind, shape)
    u = bgk.hit_velocity(*shape, 2.0, 0.03, 17)
    p = bgk.Lattice(*shape, omega=1.93, storage=kind, two_step=False)
    s = bgk.Lattice(*shape, omega=1.93, storage=kind, layout="slab", two_step="always",
                    timing=True)
    p.init_equilibrium(u)
    s.init_equilibrium(u)
    for n in (3, 1, 5, 6, 4):
        p.step(n)
        s.step(n)
        assert np.array_equal(s.get_populations(), p.get_populations()), n
    t = s.timing_passes()
    assert (t[2][1], t[3][1]) == (3, 4)


@pytest.mark.parametrize("kind", ["f64", "f32", "f16"])
@pytest.mark.parametrize("three", [True, False])
def test_slab_ghosts_after_fused_passes(gpu, kind, three):
    """The slab's reference-compatible dissipation reads the ghost planes as
    the exchanges left them (images of the state two steps back): after
    fused passes, mixed with single steps, it equals the single-step run's
    bit for bit."""
    say("slab ghosts", kind, three)
    shape = (40, 24, 16)
    nu = bgk.svisc_from_omega(1.93)
    u = bgk.hit_velocity(*shape, 2.0, 0.03, 23)
    a = bgk.Lattice(*shape, omega=1.93, storage=kind, layout="slab", two_step="always",
                    three_step=three)
    b = bgk.Lattice(*shape, omega=1.93, storage=kind, layout="slab", two_step=False)
    a.init_equilibrium(u)
    b.init_equilibrium(u)
    for n in (3, 1, 2, 1, 4, 1, 5, 2, 3):
        a.step(n)
        b.step(n)
        assert np.array_equal(a.get_populations(), b.get_populations()), n
        assert np.array_equal(a.dissipation(nu, refcompat=True),
                              b.dissipation(nu, refcompat=True)), n


@pytest.mark.parametrize("kind", KINDS3)
@pytest.mark.parametrize("layout", ["periodic", "slab"])
def test_fused_passes_against_oracle(gpu, kind, layout):
    """Three- and two-step passes (forced at a small size) against the CPU
    oracle's single steps directly: 3 + 2 + 3 + 3 steps, populations bit for
    bit."""
    import oracle

    say("fused vs oracle", kind, layout)
    shape, omega = (40, 24, 16), 1.93
    u = bgk.hit_velocity(*shape, 3.0, 0.03, 31)
    g = bgk.Lattice(*shape, omega=omega, storage=kind, layout=layout, two_step="always",
                    timing=True)
    o = oracle.OracleLattice(*shape, omega=omega, kind=kind)
    g.init_equilibrium(u)
    o.init_equilibrium(u)
    for n in (3, 2, 6):
        g.step(n)
        o.step(n)
        assert np.array_equal(g.get_populations(), o.get_populations()), n
    t = g.timing_passes()
    assert (t[2][1], t[3][1]) == (1, 3)
