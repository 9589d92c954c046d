"""Pin the CPU oracle (oracle/bgk_oracle.c) before trusting it as the checker.

The reference ships no tests, fixtures or golden vectors for this path, and
its sources cannot be built here, so each oracle function is pinned by
closed forms, invariants and an independent generic-table numpy restatement
(tests/d3q19_numpy.py).  Each check is chosen so that a plausible slip in
the oracle's hand-unrolled code (a wrong sign, neighbour index or
population, a missing term) fails it.
"""
import numpy as np
import pytest

import oracle
from tests import d3q19_numpy as ref

RNG = np.random.default_rng(20261203)


def random_pops(shape, amp=2e-3, rng=RNG):
    lx, ly, lz = shape
    return amp * rng.standard_normal((19, lz, ly, lx))


def small_velocity(shape, amp=0.03, rng=RNG):
    lx, ly, lz = shape
    return amp * rng.standard_normal((3, lz, ly, lx))


# ---------------------------------------------------------------- tables
def test_velocity_set_identities():
    """sum w = 1, sum w c = 0, sum w c c = I/3, opposite pairs exist."""
    Cf = ref.C.astype(float)
    assert abs(ref.W.sum() - 1) < 1e-15
    assert np.abs(ref.W @ Cf).max() < 1e-15
    assert np.allclose(np.einsum("q,qa,qb->ab", ref.W, Cf, Cf), np.eye(3) / 3, atol=1e-15)
    for q in range(19):
        assert ((ref.C == -ref.C[q]).all(axis=1)).sum() == 1


# ------------------------------------------------------------- streaming
@pytest.mark.parametrize("kind", ["f64", "f32"])
def test_omega_zero_is_exact_periodic_shift(kind):
    """omega = 0: b = x (pulled), so T steps move population q by T c_q
    exactly -- the semantics of the reference's streaming test kernels
    (SRC/do_somethingGPU.F90:63-73)."""
    shape = (6, 5, 7)
    f0 = random_pops(shape)
    if kind == "f32":
        f0 = f0.astype(np.float32).astype(np.float64)
    o = oracle.OracleLattice(*shape, omega=0.0, kind=kind)
    o.set_populations(f0)
    T = 3
    o.step(T)
    f = o.get_populations()
    exp = f0.copy()
    for _ in range(T):
        exp = ref.pull(exp)
    assert np.array_equal(f, exp)


def test_bcond_ghosts_are_periodic_images():
    """After SRC/bcond.comm.step3 sweeps every ghost a pull reads equals the
    periodic image of the interior value."""
    shape = (5, 4, 6)
    lx, ly, lz = shape
    f0 = random_pops(shape)
    o = oracle.OracleLattice(*shape, omega=1.0)
    o.set_populations(f0)
    # scramble halos by running one step, then refresh
    o.step(1)
    o.bcond_step3()
    f = o.get_populations()
    for q in range(18):
        h = o.halo_array(q)
        cx, cy, cz = ref.C[q]
        # every site read by the pull: (i - cx, j - cy, k - cz), interior i in 1..l
        ii = np.arange(1, lx + 1) - cx
        jj = np.arange(1, ly + 1) - cy
        kk = np.arange(1, lz + 1) - cz
        got = h[np.ix_(kk, jj, ii)]
        exp = f[q][np.ix_((kk - 1) % lz, (jj - 1) % ly, (ii - 1) % lx)]
        assert np.array_equal(got, exp), q


# ------------------------------------------------------------- collision
def test_one_step_matches_generic_restatement():
    shape = (7, 6, 5)
    f0 = random_pops(shape)
    o = oracle.OracleLattice(*shape, omega=1.7)
    o.set_populations(f0)
    o.step(1)
    got = o.get_populations()
    exp = ref.bgk_step(f0, 1.7)
    assert np.abs(got - exp).max() < 1e-16 + 1e-13 * np.abs(exp).max()


def test_omega_one_gives_equilibrium_of_pulled_moments():
    shape = (4, 5, 6)
    f0 = random_pops(shape)
    o = oracle.OracleLattice(*shape, omega=1.0)
    o.set_populations(f0)
    o.step(1)
    g = ref.pull(f0)
    rho, u = ref.moments(g)
    assert np.abs(o.get_populations() - ref.feq(rho, u)).max() < 1e-16


def test_ten_steps_match_generic_restatement():
    shape = (8, 8, 8)
    f = ref.equilibrium_pops(small_velocity(shape)) + random_pops(shape, 1e-4)
    o = oracle.OracleLattice(*shape, omega=1.9)
    o.set_populations(f)
    for _ in range(10):
        f = ref.bgk_step(f, 1.9)
    o.step(10)
    assert np.abs(o.get_populations() - f).max() < 1e-14


def test_noshift_storage_matches_shifted():
    shape = (5, 5, 5)
    fs = random_pops(shape)
    fu = fs + ref.W[:, None, None, None]
    a = oracle.OracleLattice(*shape, omega=1.8)
    b = oracle.OracleLattice(*shape, omega=1.8, shifted=False)
    a.set_populations(fs)
    b.set_populations(fu)
    a.step(4)
    b.step(4)
    d = a.get_populations() + ref.W[:, None, None, None] - b.get_populations()
    assert np.abs(d).max() < 1e-14


def test_mass_and_momentum_conserved():
    shape = (9, 7, 8)
    f0 = ref.equilibrium_pops(small_velocity(shape)) + random_pops(shape, 1e-4)
    o = oracle.OracleLattice(*shape, omega=1.95)
    o.set_populations(f0)
    m0 = f0.sum()
    j0 = np.einsum("qa,qzyx->a", ref.C.astype(float), f0)
    o.step(50)
    f = o.get_populations()
    assert abs(f.sum() - m0) < 1e-12
    assert np.abs(np.einsum("qa,qzyx->a", ref.C.astype(float), f) - j0).max() < 1e-12


def test_uniform_equilibrium_is_fixed_point():
    shape = (4, 4, 4)
    u = np.zeros((3, 4, 4, 4))
    u[0], u[1], u[2] = 0.03, -0.02, 0.01
    f0 = ref.equilibrium_pops(u)
    o = oracle.OracleLattice(*shape, omega=1.9)
    o.set_populations(f0)
    o.step(20)
    assert np.abs(o.get_populations() - f0).max() < 1e-16


@pytest.mark.parametrize("omega", [1.0, 1.6])
def test_shear_wave_decay_rate(omega):
    """u_x = U sin(2 pi z / n) decays as exp(-nu k^2 t), nu = (1/omega - 1/2)/3
    (the viscosity behind SRC/hencol.F90:31)."""
    n = 64
    shape = (2, 2, n)
    z = np.arange(n)
    u = np.zeros((3, n, 2, 2))
    u[0] = 1e-4 * np.sin(2 * np.pi * z / n)[:, None, None]
    o = oracle.OracleLattice(*shape, omega=omega)
    o.init_equilibrium(u)
    s = np.sin(2 * np.pi * z / n)

    def amp():
        rho, uu = ref.moments(o.get_populations())
        return (uu[0].mean(axis=(1, 2)) * s).sum() / (s * s).sum()

    o.step(20)
    a0 = amp()
    T = 400
    o.step(T)
    a1 = amp()
    nu = (1 / omega - 0.5) / 3
    rate = -np.log(a1 / a0) / T
    k2 = (2 * np.pi / n) ** 2
    assert abs(rate / (nu * k2) - 1) < 0.01


def test_masked_bulk_plus_border_equals_full():
    """SRC/col_MC_masked.F90 with opt 0 then opt 1 covers every site once."""
    shape = (12, 10, 11)
    f0 = random_pops(shape)
    a = oracle.OracleLattice(*shape, omega=1.9)
    b = oracle.OracleLattice(*shape, omega=1.9)
    a.set_populations(f0)
    b.set_populations(f0)
    a.bcond_step3()
    a.col_MC()
    b.bcond_step3()
    b.col_MC_masked(0, 3)
    b.col_MC_masked(1, 3)
    assert np.array_equal(a.get_populations(), b.get_populations())


# --------------------------------------------------------- initialisation
def test_init_equilibrium_moments():
    """SRC/init.F90:97-137 with crho = 1: rho = 1 and j = u up to roundoff."""
    shape = (6, 5, 4)
    u = small_velocity(shape)
    o = oracle.OracleLattice(*shape, omega=1.9)
    o.init_equilibrium(u)
    f = o.get_populations()
    assert np.abs(f - ref.equilibrium_pops(u)).max() < 1e-16  # numpy side cancels 1 - 1
    rho, uu = ref.moments(f)
    assert np.abs(rho - 1).max() < 1e-15
    assert np.abs(uu * rho - u).max() < 1e-16


# ------------------------------------------------------------ diagnostics
def test_diagno_matches_restatement_and_ztot_bug():
    shape = (6, 7, 5)
    f = ref.equilibrium_pops(small_velocity(shape)) + random_pops(shape, 1e-4)
    o = oracle.OracleLattice(*shape, omega=1.9)
    o.set_populations(f)
    rho, u = ref.moments(f)
    N = f[0].size
    exp = np.array([rho.mean(), u[0].mean(), u[1].mean(), u[2].mean(),
                    (0.5 * (u * u).sum(axis=0)).mean()])
    got = o.diagno()
    assert np.abs(got - exp).max() < 1e-15
    # SRC/diagno.F90:110: ztot = ytot + zj(l,m,n) after the loop
    bug = o.diagno(ztot_bug=True)
    assert abs(bug[3] - (u[1].sum() + u[2][-1, -1, -1]) / N) < 1e-15
    assert np.array_equal(bug[[0, 1, 2, 4]], got[[0, 1, 2, 4]])


def test_diagno_uniform_state():
    shape = (4, 4, 4)
    u = np.zeros((3, 4, 4, 4))
    u[0], u[1], u[2] = 0.01, 0.02, -0.03
    o = oracle.OracleLattice(*shape, omega=1.9)
    o.set_populations(ref.equilibrium_pops(u))
    d = o.diagno()
    assert np.allclose(d, [1, 0.01, 0.02, -0.03, 0.5 * 0.0014], atol=1e-15)


def test_diagno_accumulation_in_the_compute_kind():
    """acc_mykind: the reference's FP32 build sums diagno's per-site moments
    serially in single precision (SRC/diagno.F90:107-111).  On a uniform
    state every site contributes the same value v of each moment; the FP64
    mean of N identical floats is v itself (N v exact), and the FP32 sums
    must be the serial float32 sums of N copies of v, which numpy's
    sequential cumsum computes.  At N = 2^18 the float32 rounding is
    visible against the FP64 sums."""
    shape = (64, 64, 64)
    N = 64 ** 3
    u = np.zeros((3, 64, 64, 64))
    u[0], u[1], u[2] = 0.031, -0.017, 0.0093
    o = oracle.OracleLattice(*shape, omega=1.9, kind="f32")
    o.init_equilibrium(u)
    d64, d32 = o.diagno(), o.diagno(acc_mykind=True)
    serial = {}
    for k in range(5):
        v = np.float32(d64[k])
        assert float(v) == d64[k]
        serial[k] = np.cumsum(np.full(N, v, dtype=np.float32), dtype=np.float32)[-1]
        assert d32[k] == ((float(serial[k]) / 64) / 64) / 64, k
    assert np.abs(d32 - d64).max() > 1e-6 * np.abs(d64).max()
    # with the ztot bug the FP32 running sum ends as ytot + zj(l,m,n) in float32
    bug = o.diagno(ztot_bug=True, acc_mykind=True)
    zt = np.float32(serial[2] + np.float32(d64[3]))
    assert bug[3] == ((float(zt) / 64) / 64) / 64


def test_dissipation_zero_on_uniform_equilibrium():
    shape = (4, 5, 6)
    u = np.zeros((3, 6, 5, 4))
    u[0], u[1], u[2] = 0.02, -0.01, 0.015
    o = oracle.OracleLattice(*shape, omega=1.9)
    o.set_populations(ref.equilibrium_pops(u))
    o.bcond_step3()
    diss, tke = o.dissipation(0.01)
    assert abs(diss) < 1e-30
    assert abs(tke - 0.5 * (0.02 ** 2 + 0.01 ** 2 + 0.015 ** 2) * 120) < 1e-15


def test_dissipation_fresh_halos_matches_restatement():
    shape = (6, 5, 7)
    f = ref.equilibrium_pops(small_velocity(shape)) + random_pops(shape, 1e-4)
    o = oracle.OracleLattice(*shape, omega=1.9)
    o.set_populations(f)
    o.bcond_step3()
    nu = 0.0123
    got = o.dissipation(nu)
    exp = ref.dissipation(ref.pull(f), nu)
    assert abs(got[0] - exp[0]) < 1e-12 * abs(exp[0])
    assert abs(got[1] - exp[1]) < 1e-13 * abs(exp[1])


def test_dissipation_reads_stale_halos_of_two_steps_before():
    """The reference calls dissipation after the swap without an exchange:
    the current array's ghosts hold the images of the state two steps
    earlier (exchanged when that array was last current)."""
    shape = (6, 5, 4)
    lx, ly, lz = shape
    f0 = ref.equilibrium_pops(small_velocity(shape)) + random_pops(shape, 1e-4)
    o = oracle.OracleLattice(*shape, omega=1.9)
    o.set_populations(f0)
    states = [f0]
    for _ in range(4):
        o.step(1)
        states.append(o.get_populations())
    S_t, S_tm2 = states[4], states[2]
    # pulled populations: interior sources from S_t, out-of-box sources from S_{t-2}
    g = np.empty_like(S_t)
    z, y, x = np.meshgrid(np.arange(lz), np.arange(ly), np.arange(lx), indexing="ij")
    for q in range(19):
        cx, cy, cz = ref.C[q]
        sx, sy, sz = x - cx, y - cy, z - cz
        out = (sx < 0) | (sx >= lx) | (sy < 0) | (sy >= ly) | (sz < 0) | (sz >= lz)
        g[q] = np.where(out, S_tm2[q][sz % lz, sy % ly, sx % lx], S_t[q][sz % lz, sy % ly, sx % lx])
    nu = 0.01
    got = o.dissipation(nu)
    exp = ref.dissipation(g, nu)
    assert abs(got[0] - exp[0]) < 1e-12 * abs(exp[0])
    assert abs(got[1] - exp[1]) < 1e-13 * abs(exp[1])
    fresh = ref.dissipation(ref.pull(S_t), nu)
    assert abs(fresh[0] - exp[0]) > 1e-6 * abs(exp[0])  # the two modes differ


# -------------------------------------------------------- other kinds
@pytest.mark.parametrize("kind,tol", [("f32", 2e-5), ("f32mix", 2e-7), ("f16", 3e-3),
                                      ("bf16", 2e-2)])
def test_reduced_kinds_track_f64(kind, tol):
    shape = (8, 8, 8)
    f0 = ref.equilibrium_pops(small_velocity(shape, 0.02))
    a = oracle.OracleLattice(*shape, omega=1.9)
    b = oracle.OracleLattice(*shape, omega=1.9, kind=kind)
    a.set_populations(f0)
    b.set_populations(f0)
    a.step(5)
    b.step(5)
    fa, fb = a.get_populations(), b.get_populations()
    assert np.abs(fa - fb).max() < tol * np.abs(fa).max()


def test_f16_conversion_matches_numpy_exhaustive():
    lib = oracle.load()
    bits = np.arange(65536, dtype=np.uint16)
    vals = bits.view(np.float16).astype(np.float32)
    finite = np.isfinite(vals)
    got = np.array([lib.orc_f16_to_f32(int(b)) for b in bits[finite][::7]], dtype=np.float32)
    assert np.array_equal(got, vals[finite][::7])
    # float32 -> half round to nearest even, against numpy's conversion
    rng = np.random.default_rng(7)
    f = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 1e-3,
                        rng.standard_normal(5000).astype(np.float32) * 1e-6,
                        (vals[finite][::13] * np.float32(1.0004)),
                        np.float32([6.1e-5, 5.96e-8, 2.98e-8, 65504, 65519.9, 65520])])
    got = np.array([lib.orc_f32_to_f16(float(v)) for v in f], dtype=np.uint16)
    with np.errstate(over="ignore"):
        exp = f.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, exp)


def test_bf16_conversion_round_to_nearest_even():
    lib = oracle.load()
    rng = np.random.default_rng(8)
    f = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 1e-3,
                        np.float32([1.0, 1.00390625, 1.005859375, 1.0078125])])
    got = np.array([lib.orc_f32_to_bf16(float(v)) for v in f], dtype=np.uint32)
    # nearest representable with ties to even, by brute force over the two neighbours
    b = f.view(np.uint32)
    lo = (b >> 16) << 16
    hi = lo + 0x10000
    dlo = np.abs(f.astype(np.float64) - lo.view(np.float32).astype(np.float64))
    dhi = np.abs(f.astype(np.float64) - hi.view(np.float32).astype(np.float64))
    exp = np.where(dlo < dhi, lo >> 16, np.where(dhi < dlo, hi >> 16,
                                                 np.where(((lo >> 16) & 1) == 0, lo >> 16,
                                                          hi >> 16)))
    assert np.array_equal(got, exp)
    for v in (1.0, -2.5, 3.140625):
        assert lib.orc_bf16_to_f32(lib.orc_f32_to_bf16(v)) == np.float32(v)


# ---------------------------------------------------------------- adjoint
def test_adjoint_matches_central_differences():
    shape = (5, 4, 6)
    lx, ly, lz = shape
    omega, T = 1.7, 6
    u0 = small_velocity(shape, 0.03)
    J, g = oracle.adjoint_ke(lx, ly, lz, omega, T, u0)
    assert abs(J - oracle.forward_ke(lx, ly, lz, omega, T, u0)) < 1e-15 * abs(J)
    rng = np.random.default_rng(3)
    for _ in range(3):
        d = rng.standard_normal(u0.shape) * 0.03
        eps = 1e-4
        jp = oracle.forward_ke(lx, ly, lz, omega, T, u0 + eps * d)
        jm = oracle.forward_ke(lx, ly, lz, omega, T, u0 - eps * d)
        fd = (jp - jm) / (2 * eps)
        ad = (g * d).sum()
        assert abs(fd - ad) < 1e-7 * abs(fd)
