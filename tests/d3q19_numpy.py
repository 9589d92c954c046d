"""Independent dense numpy restatement of D3Q19 BGK (test helper).

Written in the generic table form -- velocity set and weights as arrays,
equilibrium w_i rho (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u^2), streaming by
np.roll, Pi^neq = sum_i c_i c_i (f_i - feq_i) -- instead of the reference's
hand-unrolled per-population expressions, so a wrong sign, index or
population in the oracle's unrolled code shows up as a mismatch.
Arrays are (q, z, y, x), population q = 1..19 at index q-1.
"""
import numpy as np

# D3Q19 velocities in the reference's numbering (SRC/alloca.F90:57-175,
# SRC/bgk3d_HIT.F90:23-41)
C = np.array([
    (1, -1, 0), (1, 0, -1), (1, 1, 0), (1, 0, 1), (1, 0, 0), (0, 0, 1), (0, 1, 1),
    (0, 1, 0), (0, 1, -1), (-1, -1, 0), (-1, 0, -1), (-1, 1, 0), (-1, 0, 1), (-1, 0, 0),
    (0, 0, -1), (0, -1, -1), (0, -1, 0), (0, -1, 1), (0, 0, 0)], dtype=np.int64)
W = np.array([{0: 1 / 3, 1: 1 / 18, 2: 1 / 36}[int((c * c).sum())] for c in C])


def pull(f):
    """g_q(x) = f_q(x - c_q), periodic."""
    g = np.empty_like(f)
    for q in range(19):
        cx, cy, cz = C[q]
        g[q] = np.roll(f[q], shift=(cz, cy, cx), axis=(0, 1, 2))
    return g


def moments(f, shifted=True):
    rho = f.sum(axis=0) + (1.0 if shifted else 0.0)
    j = np.einsum("qa,qzyx->azyx", C.astype(np.float64), f)
    return rho, j / rho


def feq(rho, u, shifted=True):
    cu = np.einsum("qa,azyx->qzyx", C.astype(np.float64), u)
    usq = (u * u).sum(axis=0)
    e = W[:, None, None, None] * rho[None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq[None])
    return e - (W[:, None, None, None] if shifted else 0.0)


def bgk_step(f, omega, shifted=True):
    g = pull(f)
    rho, u = moments(g, shifted)
    return g - omega * (g - feq(rho, u, shifted))


def pi_neq(g, shifted=True):
    rho, u = moments(g, shifted)
    n = g - feq(rho, u, shifted)
    Cf = C.astype(np.float64)
    return np.einsum("qa,qb,qzyx->abzyx", Cf, Cf, n), rho, u


def dissipation(g, svisc, shifted=True):
    """(sum 2 nu Pi:Pi, sum |u|^2/2) from already pulled populations g."""
    P, rho, u = pi_neq(g, shifted)
    return (2 * svisc * (P * P).sum(axis=(0, 1))).sum(), (0.5 * (u * u).sum(axis=0)).sum()


def equilibrium_pops(u):
    """Shifted populations of the equilibrium of u with rho = 1."""
    return feq(np.ones(u.shape[1:]), u)
