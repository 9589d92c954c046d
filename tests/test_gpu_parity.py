"""GPU parity: the HIP library (through the C-ABI) against the CPU oracle.

Bar: FP64/FP32/F32-mixed/F16/BF16 populations bit-identical to the oracle's
restatement of the reference arithmetic (same IEEE operations, no FMA
contraction on either side); diagnostics within 1e-10 relative in FP64 (the
GPU sums in a tree, the oracle serially).  At full sizes the checks are
size-independent properties (conservation, slab/periodic identity).
"""
import sys

import numpy as np
import pytest

import bgk3d_hit_amd as bgk
import oracle
from tests import d3q19_numpy as ref

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def hit_pair(shape, k0, urms, seed, omega, kind="f64", refcompat=False):
    u = bgk.hit_velocity(*shape, k0, urms, seed)
    g = bgk.Lattice(*shape, omega=omega, storage=kind, refcompat_halos=refcompat)
    o = oracle.OracleLattice(*shape, omega=omega, kind=kind)
    g.init_equilibrium(u)
    o.init_equilibrium(u)
    return g, o, u


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix", "f16", "bf16"])
def test_init_equilibrium_bitwise(gpu, kind):
    shape = (12, 10, 8)
    say("init_equilibrium", kind, shape)
    g, o, _ = hit_pair(shape, 2.0, 0.03, 7, 1.9, kind)
    assert np.array_equal(g.get_populations(), o.get_populations())


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix", "f16", "bf16"])
def test_one_step_bitwise(gpu, kind):
    shape = (70, 9, 6)  # x not a multiple of the 64-wide tile
    say("one step", kind, shape)
    rng = np.random.default_rng(1)
    f0 = 2e-3 * rng.standard_normal((19, shape[2], shape[1], shape[0]))
    g = bgk.Lattice(*shape, omega=1.93, storage=kind)
    o = oracle.OracleLattice(*shape, omega=1.93, kind=kind)
    g.set_populations(f0)
    o.set_populations(f0)
    assert np.array_equal(g.get_populations(), o.get_populations())
    g.step(1)
    o.step(1)
    assert np.array_equal(g.get_populations(), o.get_populations())


def test_pure_streaming_shift(gpu):
    """omega = 0: exact periodic shift by c_q per step
    (SRC/do_somethingGPU.F90:63-73 semantics)."""
    shape = (65, 7, 5)
    say("pure streaming", shape)
    rng = np.random.default_rng(2)
    f0 = rng.standard_normal((19, 5, 7, 65))
    g = bgk.Lattice(*shape, omega=0.0)
    g.set_populations(f0)
    g.step(4)
    exp = f0
    for _ in range(4):
        exp = ref.pull(exp)
    assert np.array_equal(g.get_populations(), exp)


@pytest.mark.parametrize("kind,nsteps", [("f64", 200), ("f32", 200), ("f32mix", 100),
                                         ("f16", 100), ("bf16", 100)])
def test_hit_trajectory_bitwise(gpu, kind, nsteps):
    shape = (32, 32, 32)
    say("trajectory", kind, shape, nsteps)
    g, o, _ = hit_pair(shape, 4.0, 0.02, 1234, 1.90, kind)
    g.step(nsteps)
    o.step(nsteps)
    assert np.array_equal(g.get_populations(), o.get_populations())


def test_hit_64cubed_fp64_1000_steps_with_diagnostics(gpu):
    """64^3 HIT, k0=4, urms=0.02, omega=1.90, FP64, 1000 steps, diagnostics
    every 100 steps in the reference loop order."""
    shape = (64, 64, 64)
    omega = 1.90
    nu = bgk.svisc_from_omega(omega)
    say("64^3 x 1000 FP64")
    g, o, _ = hit_pair(shape, 4.0, 0.02, 1234, omega, "f64", refcompat=True)
    for it in range(1, 1001):
        g.boundaries()
        g.col_MC()
        o.bcond_step3()
        o.col_MC()
        if it % 100 == 0:
            dg, do = g.diagno(), o.diagno()
            assert abs(dg[0] - do[0]) < 1e-12  # oracle sums 64^3 values serially
            assert np.abs(dg[1:4] - do[1:4]).max() < 1e-10 * 0.02
            assert abs(dg[4] - do[4]) < 1e-10 * abs(do[4])
            # reference-compatible dissipation (stale halos), then consistent
            sg, so = g.dissipation(nu, refcompat=True), o.dissipation(nu)
            assert np.abs(sg - so).max() < 1e-10 * np.abs(so).max(), (it, sg, so)
            bg = g.diagno(refcompat=True)
            bo = o.diagno(ztot_bug=True)
            assert abs(bg[3] - bo[3]) < 1e-10 * 0.02
            o.bcond_step3()
            cg, co = g.dissipation(nu), o.dissipation(nu)
            assert np.abs(cg - co).max() < 1e-10 * np.abs(co).max()
    fg, fo = g.get_populations(), o.get_populations()
    assert rel(fg, fo) <= 1e-12
    assert np.array_equal(fg, fo)


@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_masked_bulk_then_border_equals_full(gpu, kind):
    shape = (40, 36, 30)
    say("masked", kind)
    g1, _, u = hit_pair(shape, 3.0, 0.03, 3, 1.95, kind)
    g2 = bgk.Lattice(*shape, omega=1.95, storage=kind)
    g2.init_equilibrium(u)
    g2.set_border(5)
    for _ in range(3):
        g1.col_MC()
        g2.col_MC_masked(0)
        g2.col_MC_masked(1)
    assert np.array_equal(g1.get_populations(), g2.get_populations())


@pytest.mark.parametrize("kind", ["f32", "f64", "f16"])
@pytest.mark.parametrize("two_step", ["always", False])
def test_slab_overlap_bitwise_equals_periodic(gpu, kind, two_step):
    """One rank's slab with +-z ghost planes, pack / self-exchange / unpack,
    bulk and boundary collides on two streams (one step per pass, or the
    two-step pass on the bulk with the boundary planes stepped twice through
    the plane buffer): identical to the periodic kernel on the same domain."""
    shape = (64, 48, 16)
    say("slab", kind, shape, two_step)
    u = bgk.hit_velocity(*shape, 4.0, 0.03, 5678)
    p = bgk.Lattice(*shape, omega=1.97, storage=kind, two_step=False)
    s = bgk.Lattice(*shape, omega=1.97, storage=kind, layout="slab", two_step=two_step)
    s2 = bgk.Lattice(*shape, omega=1.97, storage=kind, layout="slab", two_step=two_step)
    for x in (p, s, s2):
        x.init_equilibrium(u)
    p.step(25)
    s.step(25)                       # overlapped two-stream path
    for _ in range(25):              # reference call order, one stream
        s2.boundaries()
        s2.col_MC()
    fp = p.get_populations()
    assert np.array_equal(s.get_populations(), fp)
    assert np.array_equal(s2.get_populations(), fp)
    # consistent dissipation refreshes the ghosts first: same as periodic
    nu = bgk.svisc_from_omega(1.97)
    assert np.allclose(s.dissipation(nu), p.dissipation(nu), rtol=1e-13, atol=0)


@pytest.mark.parametrize("kind", ["f32", "f32mix", "bf16"])
@pytest.mark.parametrize("shape", [(70, 9, 6), (33, 20, 7), (64, 16, 40)])
def test_slab_two_step_ragged(gpu, kind, shape):
    """The slab's two-step path at the smallest slab it takes (6 planes),
    odd plane counts and ragged x/y, odd and even step counts."""
    say("slab two-step ragged", kind, shape)
    u = bgk.hit_velocity(*shape, 2.0, 0.03, 13)
    p = bgk.Lattice(*shape, omega=1.93, storage=kind, two_step=False)
    s = bgk.Lattice(*shape, omega=1.93, storage=kind, layout="slab", two_step="always",
                    timing=True, three_step=False)
    p.init_equilibrium(u)
    s.init_equilibrium(u)
    for n in (2, 3, 1, 6):
        p.step(n)
        s.step(n)
        assert np.array_equal(s.get_populations(), p.get_populations()), n
    assert s.timing()["collide_launches"] < 12 * 3  # took two-step passes


def test_slab_masked_path(gpu):
    shape = (32, 32, 12)
    say("slab masked")
    u = bgk.hit_velocity(*shape, 2.0, 0.03, 9)
    p = bgk.Lattice(*shape, omega=1.9, storage="f32")
    s = bgk.Lattice(*shape, omega=1.9, storage="f32", layout="slab")
    s.set_border(2)
    p.init_equilibrium(u)
    s.init_equilibrium(u)
    for _ in range(5):
        p.step(1)
        s.boundaries()
        s.col_MC_masked(0)
        s.col_MC_masked(1)
    assert np.array_equal(s.get_populations(), p.get_populations())


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix", "f16", "bf16"])
@pytest.mark.parametrize("layout", ["periodic", "slab"])
@pytest.mark.parametrize("shape,border", [((23, 17, 13), 3), ((20, 14, 11), 1),
                                          ((16, 12, 9), 4), ((12, 10, 9), 5)])
def test_masked_regions_against_oracle(gpu, kind, layout, shape, border):
    """col_MC_masked against the oracle's restatement of the reference mask
    (SRC/init.F90:44-62; SRC/col_MC_masked.F90:114,355-369).  The bulk call
    (opt 0) runs at one omega and the border call (opt 1) at another, so
    every site's result shows which call updated it: a swapped, missing or
    doubled region changes the bits.  (12, 10, 9) with border 5 has no bulk."""
    say("masked vs oracle", kind, layout, shape, border)
    w_bulk, w_border = 1.93, 1.71
    g, o, _ = hit_pair(shape, 2.0, 0.03, 31, w_bulk, kind)
    if layout == "slab":
        g.close()
        g = bgk.Lattice(*shape, omega=w_bulk, storage=kind, layout="slab")
        g.init_equilibrium(bgk.hit_velocity(*shape, 2.0, 0.03, 31))
    g.set_border(border)
    for _ in range(3):
        if layout == "slab":
            g.boundaries()
        g.set_omega(w_bulk)
        g.col_MC_masked(0)
        g.set_omega(w_border)
        g.col_MC_masked(1)
        o.bcond_step3()
        o.set_omega(w_bulk)
        o.col_MC_masked(0, border)
        o.set_omega(w_border)
        o.col_MC_masked(1, border)
    fg, fo = g.get_populations(), o.get_populations()
    assert np.array_equal(fg, fo), np.abs(fg - fo).max()
    # the two omegas must leave different results (the test would pin
    # nothing otherwise): a full step at the border omega differs
    if 2 * border < min(shape):
        f_all = bgk.Lattice(*shape, omega=w_border, storage=kind)
        f_all.init_equilibrium(bgk.hit_velocity(*shape, 2.0, 0.03, 31))
        f_all.step(3)
        assert not np.array_equal(f_all.get_populations(), fg)


def test_device_hit_equals_host_hit(gpu):
    shape = (40, 32, 24)
    say("device HIT")
    u = bgk.hit_velocity(*shape, 3.0, 0.025, 42)
    a = bgk.Lattice(*shape, omega=1.9)
    b = bgk.Lattice(*shape, omega=1.9)
    a.init_equilibrium(u)
    b.init_hit(3.0, 0.025, 42)
    assert np.array_equal(a.get_populations(), b.get_populations())


def test_nonfinite_populations_are_reported(gpu):
    shape = (16, 16, 16)
    say("nan injection")
    g = bgk.Lattice(*shape, omega=1.9)
    f = np.zeros((19, 16, 16, 16))
    f[3, 5, 6, 7] = np.nan
    g.set_populations(f)
    g.step(2)
    with pytest.raises(bgk.BGKError) as e:
        g.diagno()
    assert e.value.code == 5
    with pytest.raises(bgk.BGKError):
        g.dissipation(0.01)


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix", "f16", "bf16"])
@pytest.mark.parametrize("three_step", [True, False])
@pytest.mark.parametrize("shape", [(70, 9, 20), (33, 21, 7), (130, 24, 36)])
def test_refcompat_images_with_fused_passes(gpu, kind, three_step, shape):
    """With the reference's stale halo images a two-step pass records the
    faces of its source state and writes those of its intermediate state; a
    three-step pass writes the faces of both of its intermediate states.  Only
    the passes that end at the call's last step or the one before write
    them (7, 8, 10, 11 steps: passes without images first).  The refcompat
    dissipation after any mix of passes and single steps equals the
    single-step path's exactly."""
    say("refcompat fused", kind, three_step, shape)
    u = bgk.hit_velocity(*shape, 2.0, 0.03, 17)
    a = bgk.Lattice(*shape, omega=1.93, storage=kind, refcompat_halos=True, two_step="always",
                    three_step=three_step, timing=True)
    b = bgk.Lattice(*shape, omega=1.93, storage=kind, refcompat_halos=True, two_step=False)
    a.init_equilibrium(u)
    b.init_equilibrium(u)
    nu = bgk.svisc_from_omega(1.93)
    for n in (2, 3, 1, 2, 4, 5, 0, 6, 2, 9, 7, 1, 8, 0, 10, 1, 1, 11):
        if n == 0:  # the reference call order, one step
            for x in (a, b):
                x.boundaries()
                x.col_MC()
        else:
            a.step(n)
            b.step(n)
        assert np.array_equal(a.get_populations(), b.get_populations()), n
        assert np.array_equal(a.dissipation(nu, refcompat=True),
                              b.dissipation(nu, refcompat=True)), n
        assert np.array_equal(a.diagnostics(nu, refcompat=True)[1],
                              b.dissipation(nu, refcompat=True)), n
    depth = a.timing_passes()
    assert (depth[3][1] > 0) == three_step, depth
    assert depth[2][1] > 0, depth


def test_refcompat_images_512cubed_whole_columns(gpu):
    """512^3 FP32 with reference halo images: six steps are two three-step
    passes over whole tile columns and chunks, the second writing the face
    images; the refcompat dissipation and diagnostics then equal those of
    the reference call order (boundaries + col_MC every step), as do the
    populations (checksums)."""
    say("refcompat 512^3 whole columns")
    n, omega = 512, 1.95
    nu = bgk.svisc_from_omega(omega)
    a = bgk.Lattice(n, n, n, omega=omega, storage="f32", refcompat_halos=True, timing=True)
    a.init_hit(4.0, 0.02, 5)
    a.step(6)
    assert a.timing_passes()[3][1] == 2
    da = a.dissipation(nu, refcompat=True)
    ga = a.diagno(refcompat=True)
    ca = a.checksum()
    a.close()
    b = bgk.Lattice(n, n, n, omega=omega, storage="f32", refcompat_halos=True, two_step=False)
    b.init_hit(4.0, 0.02, 5)
    for _ in range(6):
        b.boundaries()
        b.col_MC()
    assert b.checksum() == ca
    assert np.array_equal(b.dissipation(nu, refcompat=True), da)
    assert np.array_equal(b.diagno(refcompat=True), ga)
    b.close()


def test_refcompat_dissipation_needs_flag(gpu):
    g = bgk.Lattice(8, 8, 8, omega=1.9)
    with pytest.raises(bgk.BGKError) as e:
        g.dissipation(0.01, refcompat=True)
    assert e.value.code == 4


@pytest.mark.parametrize("two_step,launches", [("always", 5), (True, 10), (False, 10)])
def test_timing_counters(gpu, two_step, launches):
    g = bgk.Lattice(64, 64, 64, omega=1.9, storage="f32", timing=True, two_step=two_step,
                    three_step=False)
    g.init_hit(4.0, 0.02, 1)
    g.step(10)
    g.diagno()
    t = g.timing()
    assert t["collide_launches"] == launches and t["reduce_launches"] == 1
    assert t["collide_ms"] > 0 and t["step_ms"] >= t["collide_ms"] * 0.5


@pytest.mark.parametrize("n,launches", [(64, 20), (128, 10), (256, 10)])
def test_two_step_engages_where_it_pays(gpu, n, launches):
    """bgk_step's chunk model: 64^3 (104 workgroups, fewer than half the
    CUs) takes single steps; 128^3 and up take two steps per pass."""
    g = bgk.Lattice(n, n, n, omega=1.9, storage="f32", timing=True, three_step=False)
    g.step(20)
    assert g.timing()["collide_launches"] == launches


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix", "f16", "bf16"])
@pytest.mark.parametrize("shape", [(70, 9, 6), (64, 16, 130), (24, 20, 16), (5, 3, 1),
                                   (130, 17, 65), (1, 1, 2)])
def test_two_step_pass_equals_single_steps(gpu, kind, shape):
    """bgk_step fuses two time steps per pass over the populations (grown
    tiles, first step kept in LDS); it must give the single-step result bit
    for bit, at sizes that leave partial tiles and a partial last z chunk,
    and for odd step counts (a fused pass followed by one single step)."""
    say("two-step vs single", kind, shape)
    u = bgk.hit_velocity(*shape, 2.0, 0.03, 11)
    a = bgk.Lattice(*shape, omega=1.93, storage=kind, two_step="always", three_step=False)
    b = bgk.Lattice(*shape, omega=1.93, storage=kind, two_step=False)
    a.init_equilibrium(u)
    b.init_equilibrium(u)
    for n in (2, 3, 1, 4):
        a.step(n)
        for _ in range(n):
            b.col_MC()
        assert np.array_equal(a.get_populations(), b.get_populations()), n


def test_two_step_pass_512cubed_fp32_checksum(gpu):
    """The headline lattice: 512^3 FP32, 10 steps with the two-step pass ==
    10 single steps, bit for bit."""
    say("two-step 512^3 f32")
    a = bgk.Lattice(512, 512, 512, omega=1.95, storage="f32", timing=True, three_step=False)
    a.init_hit(4.0, 0.02, 1234)
    a.step(10)
    ca = a.checksum()
    assert a.timing()["collide_launches"] == 5
    a.close()
    b = bgk.Lattice(512, 512, 512, omega=1.95, storage="f32", two_step=False)
    b.init_hit(4.0, 0.02, 1234)
    b.step(10)
    assert b.checksum() == ca


def test_two_step_pass_1040cubed(gpu):
    """1040^3 FP32 (84 GB of populations, tiles not dividing the lattice):
    three-step passes (3 + 3), two-step passes (2 + 2 + 2) and single steps
    give the same populations."""
    n = 1040
    sums = []
    for name, kw, launches in (("three-step", {}, 2), ("two-step", {"three_step": False}, 3),
                               ("single", {"two_step": False}, 6)):
        say(name, "1040^3 f32")
        a = bgk.Lattice(n, n, n, omega=1.95, storage="f32", timing=True, **kw)
        a.init_hit(4.0, 0.02, 77)
        a.step(6)
        sums.append(a.checksum())
        assert a.timing()["collide_launches"] == launches, name
        a.close()
    assert sums[0] == sums[1] == sums[2]


@pytest.mark.parametrize("kind", ["f32", "f16", "f64"])
def test_two_step_pass_256cubed_checksum(gpu, kind):
    """At 256^3 step() takes the two-step pass by itself (10 launches for 20
    steps) and matches single steps bit for bit."""
    say("two-step 256^3", kind)
    u = bgk.hit_velocity(256, 256, 256, 4.0, 0.02, 5)
    a = bgk.Lattice(256, 256, 256, omega=1.95, storage=kind, timing=True, three_step=False)
    b = bgk.Lattice(256, 256, 256, omega=1.95, storage=kind, two_step=False)
    a.init_equilibrium(u)
    b.init_equilibrium(u)
    a.step(20)
    b.step(20)
    assert a.timing()["collide_launches"] == 10
    assert a.checksum() == b.checksum()


def test_512_cubed_fp32_against_fp64(gpu):
    """Full-size check of the 512^3 workload: FP64 conserves mass to
    roundoff over 20 steps, and the FP32 run (the reference's
    single-precision arithmetic, bit-exact with the oracle at small sizes)
    tracks it: kinetic energy within 1e-5 relative, mean density within
    1e-6."""
    say("512^3 f32 vs f64")
    res = {}
    for kind in ("f64", "f32"):
        g = bgk.Lattice(512, 512, 512, omega=1.95, storage=kind)
        g.init_hit(4.0, 0.02, 1234)
        d0 = g.diagno()
        g.step(20)
        res[kind] = (d0, g.diagno(), g.dissipation(bgk.svisc_from_omega(1.95)))
        g.close()
    d0, d1, s64 = res["f64"]
    assert abs(d1[0] - d0[0]) < 1e-13  # mass
    # (the mean of u = j/rho is not an invariant -- sum j is -- so mean u
    # may drift physically; it is compared between the two precisions)
    assert d1[4] < d0[4]  # decaying turbulence loses kinetic energy
    e0, e1, s32 = res["f32"]
    assert abs(e1[0] - d1[0]) < 1e-6
    assert np.abs(e1[1:4] - d1[1:4]).max() < 1e-5 * 0.02
    assert abs(e1[4] - d1[4]) < 1e-5 * d1[4]
    assert np.abs(s32 - s64).max() < 1e-4 * np.abs(s64).max()


def test_checksum_sees_one_ulp(gpu):
    g = bgk.Lattice(32, 32, 32, omega=1.9, storage="f32")
    g.init_hit(4.0, 0.02, 7)
    c0 = g.checksum()
    assert c0 == g.checksum()
    f = g.get_populations()
    v = np.float32(f[3, 4, 5, 6])
    f2 = f.copy()
    f2[3, 4, 5, 6] = np.nextafter(v, np.float32(1))
    g.set_populations(f2)
    assert g.checksum() != c0
    g.set_populations(f)
    assert g.checksum() == c0


def test_hit_128cubed_fp64_against_oracle(gpu):
    """128^3 HIT, omega=1.95, FP64: GPU against the CPU oracle, diagnostics
    every 50 steps."""
    shape, omega = (128, 128, 128), 1.95
    nu = bgk.svisc_from_omega(omega)
    say("128^3 FP64 vs oracle")
    g, o, _ = hit_pair(shape, 4.0, 0.02, 1234, omega, "f64")
    for it in range(50, 201, 50):
        g.step(50)
        o.step(50)
        dg, do = g.diagno(), o.diagno()
        assert abs(dg[0] - do[0]) < 1e-12
        assert abs(dg[4] - do[4]) < 1e-10 * do[4]
        o.bcond_step3()
        sg, so = g.dissipation(nu), o.dissipation(nu)
        assert np.abs(sg - so).max() < 1e-10 * np.abs(so).max()
    assert np.array_equal(g.get_populations(), o.get_populations())


def test_slab_1024x1024x256_bitwise(gpu):
    """One rank's share at its real shape: 1024x1024x256 FP32 slab with
    +-z ghost planes and the two-stream overlap (three-step passes), over
    the full 1000 steps of the workload (k0=8, u_rms=0.03, omega=1.97):
    diagno every 50 steps and the final populations bit-identical to the
    periodic lattice."""
    shape = (1024, 1024, 256)
    say("slab 1024x1024x256 x 1000")
    p = bgk.Lattice(*shape, omega=1.97, storage="f32")
    p.init_hit(8.0, 0.03, 5678)
    dp_ = []
    for _ in range(20):
        p.step(50)
        dp_.append(p.diagno())
    cp = p.checksum()
    p.close()
    s = bgk.Lattice(*shape, omega=1.97, storage="f32", layout="slab")
    s.init_hit(8.0, 0.03, 5678)
    for k in range(20):
        s.step(50)
        assert np.array_equal(s.diagno(), dp_[k]), (k + 1) * 50
    assert s.checksum() == cp


def test_get_box_matches_full_download(gpu):
    shape = (40, 36, 30)
    g = bgk.Lattice(*shape, omega=1.9, storage="f16")
    g.init_hit(3.0, 0.03, 11)
    g.step(3)
    f = g.get_populations()
    for (x0, x1, y0, y1, z0, z1) in [(5, 6, 7, 8, 9, 10), (0, 40, 0, 36, 15, 16),
                                     (3, 17, 0, 36, 2, 29), (0, 40, 18, 19, 0, 30)]:
        assert np.array_equal(g.get_box(x0, x1, y0, y1, z0, z1), f[:, z0:z1, y0:y1, x0:x1])
    with pytest.raises(bgk.BGKError):
        g.get_box(0, 41, 0, 1, 0, 1)


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix", "f16", "bf16"])
@pytest.mark.parametrize("layout,refcompat", [("periodic", False), ("periodic", True),
                                              ("slab", False), ("slab", True)])
def test_fused_diagnostics_equal_separate_calls(gpu, kind, layout, refcompat):
    """bgk_diagnostics (diagno + dissipation in one pass) returns exactly what
    bgk_diagno and bgk_dissipation return, in every mode and layout."""
    say("fused diagnostics", kind, layout, refcompat)
    shape = (36, 22, 19)
    kw = dict(omega=1.93, storage=kind, layout=layout)
    if layout == "periodic":
        kw["refcompat_halos"] = refcompat
    g = bgk.Lattice(*shape, **kw)
    g.init_hit(4.0, 0.03, 77)
    nu = bgk.svisc_from_omega(1.93)
    for n in (0, 1, 6):
        g.step(n)
        d, s = g.diagnostics(nu, refcompat=refcompat)
        d2 = g.diagno(refcompat=refcompat)
        s2 = g.dissipation(nu, refcompat=refcompat)
        assert np.array_equal(d, d2), (n, d, d2)
        assert np.array_equal(s, s2), (n, s, s2)


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix", "f16", "bf16"])
@pytest.mark.parametrize("layout", ["periodic", "slab"])
def test_wide_offsets_equal_32bit_offsets(gpu, kind, layout):
    """The 64-bit-offset kernels (taken for x-y planes above 4 GiB, forced
    here by BGK_FLAG_WIDE_OFFSETS) give the same populations as the 32-bit
    ones, for single steps and two-step passes."""
    say("wide offsets", kind, layout)
    shape = (40, 26, 21)
    u = bgk.hit_velocity(*shape, 4.0, 0.03, 19)
    for two_step in ("always", False):
        a = bgk.Lattice(*shape, omega=1.93, storage=kind, layout=layout, two_step=two_step)
        b = bgk.Lattice(*shape, omega=1.93, storage=kind, layout=layout, two_step=two_step,
                        wide_offsets=True)
        a.init_equilibrium(u)
        b.init_equilibrium(u)
        a.step(7)
        b.step(7)
        assert np.array_equal(a.get_populations(), b.get_populations()), two_step
        a.close()
        b.close()
