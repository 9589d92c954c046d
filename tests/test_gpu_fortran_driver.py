"""The Fortran host driver (bgk3d_hit_amd/bin/bgk3d_HIT.x) end to end on the
GPU: bgk.input in, diagno/dissipation files out, checked against the oracle
run in the reference's loop order with the reference's diagnostics
(stale halos, ztot column)."""
import os
import subprocess

import numpy as np
import pytest

import bgk3d_hit_amd as bgk
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "bgk3d_hit_amd", "bin", "bgk3d_HIT.x")


def run_driver(tmp_path, deck):
    (tmp_path / "bgk.input").write_text(deck)
    p = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stdout + p.stderr
    return p


def test_driver_exists_and_needs_a_deck(tmp_path):
    assert os.path.exists(DRIVER), "build the driver with make -C bgk3d_hit_amd"
    p = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert p.returncode != 0
    assert "cannot open" in p.stdout


@pytest.mark.gpu
def test_driver_hit_f64_matches_oracle(gpu, tmp_path):
    n, svisc, itfin, icheck = 32, 0.008, 200, 50
    deck = (f"&parameters\nlx={n}\nly={n}\nlz={n}\nsvisc={svisc}\nitfin={itfin}\n"
            f"icheck={icheck}\nisignal=100\ninit='hit'\nk0=4.0\nurms=0.02\nseed=1234\n"
            f"storage='f64'\ndiag='refcompat'\n/\n")
    print("[gpu-case] fortran driver hit f64", flush=True)
    run_driver(tmp_path, deck)
    dg = np.loadtxt(tmp_path / "diagno.full.dat", ndmin=2)
    ds = np.loadtxt(tmp_path / "dissipation.full.dat", ndmin=2)
    assert (tmp_path / "diagno.dat").exists() and (tmp_path / "bgk.perf").exists()
    omega = 2.0 / (6.0 * svisc + 1.0)
    u = bgk.hit_velocity(n, n, n, 4.0, 0.02, 1234)
    o = oracle.OracleLattice(n, n, n, omega=omega, kind="f64")
    o.init_equilibrium(u)
    row = 0
    for it in range(1, itfin + 1):
        o.bcond_step3()
        o.col_MC()
        if it % icheck == 0:
            d = o.diagno(ztot_bug=True)
            s = o.dissipation(svisc)
            assert dg[row, 0] == it and ds[row, 0] == it
            # diagno.dat column order: itime, ux, uy, uz(ref), rho, ke
            assert abs(dg[row, 4] - d[0]) < 1e-14
            assert np.abs(dg[row, 1:4] - d[1:4]).max() < 1e-10 * 0.02
            assert abs(dg[row, 5] - d[4]) < 1e-10 * d[4]
            assert abs(ds[row, 2] - s[1]) < 1e-10 * s[1]
            assert abs(ds[row, 3] - s[0]) < 1e-10 * s[0]
            row += 1
    assert row == itfin // icheck


@pytest.mark.gpu
def test_driver_tgv_reference_deck_shape(gpu, tmp_path):
    """The reference's own deck shape (RUN/bgk.input: TGV), shortened."""
    deck = ("&parameters\nlx = 48\nly = 48\nlz = 48\nproc_x = 1\nproc_y = 1\nproc_z = 1\n"
            "u0 = 0.1\nsvisc = 0.0014317\nitfin = 170\nitsave = 2812\nisignal = 85\n"
            "ivtim = 2812\nicheck = 85 /\n")
    print("[gpu-case] fortran driver tgv", flush=True)
    p = run_driver(tmp_path, deck)
    assert "GLUPs" in p.stdout
    d = np.loadtxt(tmp_path / "diagno.full.dat", ndmin=2)
    assert d.shape[0] == 2
    assert np.abs(d[:, 4] - 1).max() < 1e-5          # mean rho (FP32 build)
    assert (d[1, 5] < d[0, 5]) and d[1, 5] > 0         # decaying kinetic energy


def read_vtk(path):
    raw = open(path, "rb").read()
    marker = b"VECTORS velocity float\n"  # (A23): right-justified, as the reference writes it
    i = raw.index(marker) + len(marker)
    header = raw[:i].decode().splitlines()
    data = np.frombuffer(raw[i:], dtype=">f4").reshape(-1, 3)
    return header, data


@pytest.mark.gpu
def test_driver_probe_vtk_and_restart(gpu, tmp_path):
    """probe_global (fort.88) and the VTK slices against the oracle state, and
    restart: 20 steps + restart + 20 steps == 40 steps, bit for bit."""
    n, svisc = 24, 0.01
    base = (f"lx={n}\nly={n}\nlz={n}\nsvisc={svisc}\ninit='hit'\nk0=3.0\nurms=0.03\n"
            f"seed=7\nstorage='f64'\ndiag='consistent'\n")
    print("[gpu-case] fortran driver probe/vtk/restart", flush=True)
    a = tmp_path / "a"
    a.mkdir()
    run_driver(a, f"&parameters\n{base}itfin=40\nicheck=20\nivtim=40\nitsave=20\n/\n")
    b = tmp_path / "b"
    b.mkdir()
    run_driver(b, f"&parameters\n{base}itfin=20\nicheck=20\nitsave=20\n/\n")
    run_driver(b, f"&parameters\n{base}itfin=40\nicheck=20\nitsave=20\nirestart=1\n/\n")
    ra, rb = (a / "restart.bgk").read_bytes(), (b / "restart.bgk").read_bytes()
    assert ra[:8] == b"BGK3DRST" and ra == rb
    # oracle state at step 40
    omega = 2.0 / (6.0 * svisc + 1.0)
    u = bgk.hit_velocity(n, n, n, 3.0, 0.03, 7)
    o = oracle.OracleLattice(n, n, n, omega=omega, kind="f64")
    o.init_equilibrium(u)
    o.step(40)
    f = o.get_populations()
    assert np.array_equal(np.frombuffer(ra[28:], dtype=np.float64).reshape(f.shape), f)
    # fort.88: itime, then the 19 populations of site (lz/2, ly/2, lz/2) (1-based)
    vals = (a / "fort.88").read_text().split()
    i = len(vals) - 19
    assert int(vals[i - 1]) == 40
    site = f[:, n // 2 - 1, n // 2 - 1, n // 2 - 1]
    assert np.allclose(np.array(vals[i:], dtype=float), site, rtol=1e-14, atol=0)
    # xy slice at z = n/2 (1-based): float32 of (j/rho) from double sums
    header, data = read_vtk(a / "tec_xy.00000.00000040.vtk")
    assert header[4].split() == ["DIMENSIONS", str(n), str(n), "1"]
    g = f[:, n // 2 - 1]
    C = np.array([(1, -1, 0), (1, 0, -1), (1, 1, 0), (1, 0, 1), (1, 0, 0), (0, 0, 1), (0, 1, 1),
                  (0, 1, 0), (0, 1, -1), (-1, -1, 0), (-1, 0, -1), (-1, 1, 0), (-1, 0, 1),
                  (-1, 0, 0), (0, 0, -1), (0, -1, -1), (0, -1, 0), (0, -1, 1), (0, 0, 0)])
    rho = g.sum(axis=0) + 1.0
    j = np.einsum("qa,qyx->ayx", C.astype(float), g)
    exp = (j.astype(np.float32) / rho.astype(np.float32)).reshape(3, -1).T
    assert np.allclose(data, exp, rtol=2e-6, atol=1e-9)
    assert (a / "tec_xz.00000.00000040.vtk").exists() and (a / "tec_yz.00000.00000000.vtk").exists()


@pytest.mark.gpu
def test_driver_two_step_pairs_match_single_steps(gpu, tmp_path):
    """Between output steps the driver hands the library two steps per call
    (bgk_step(2): one two-step pass on a 256^3 lattice); the diagnostics at
    every icheck equal those of single steps (bgk_col_MC) exactly."""
    n, omega, itfin, icheck = 256, 1.95, 21, 10
    svisc = (2.0 / omega - 1.0) / 6.0  # the driver's hencol with omega given
    deck = (f"&parameters\nlx={n}\nly={n}\nlz={n}\nomega={omega}\nitfin={itfin}\n"
            f"icheck={icheck}\ninit='hit'\nk0=4.0\nurms=0.02\nseed=1234\n"
            f"storage='f32'\ndiag='consistent'\n/\n")
    print("[gpu-case] fortran driver 256^3 f32 two-step pairs", flush=True)
    run_driver(tmp_path, deck)
    dg = np.loadtxt(tmp_path / "diagno.full.dat", ndmin=2)
    ds = np.loadtxt(tmp_path / "dissipation.full.dat", ndmin=2)
    lat = bgk.Lattice(n, n, n, omega=omega, storage="f32", two_step=False)
    lat.init_hit(4.0, 0.02, 1234)
    for row, it in enumerate(range(icheck, itfin + 1, icheck)):
        lat.step(icheck)
        d = lat.diagno()
        s = lat.dissipation(svisc)
        assert dg[row, 0] == it and ds[row, 0] == it
        assert np.array_equal(dg[row, 1:], [d[1], d[2], d[3], d[0], d[4]]), (it, dg[row], d)
        assert ds[row, 2] == s[1] and ds[row, 3] == s[0], (it, ds[row], s)
    assert dg.shape[0] == itfin // icheck


@pytest.mark.gpu
def test_driver_refcompat_fused_passes(gpu, tmp_path):
    """The reference-compatible diagnostics (stale halo images) with the
    driver's fused passes (7 steps between outputs: a three-step pass and two
    two-step passes) on a 256^3 lattice: equal to the reference call order
    (boundaries + col_MC every step) exactly."""
    n, omega, itfin, icheck = 256, 1.95, 21, 7
    svisc = (2.0 / omega - 1.0) / 6.0
    deck = (f"&parameters\nlx={n}\nly={n}\nlz={n}\nomega={omega}\nitfin={itfin}\n"
            f"icheck={icheck}\ninit='hit'\nk0=4.0\nurms=0.02\nseed=99\n"
            f"storage='f32'\ndiag='refcompat'\n/\n")
    print("[gpu-case] fortran driver 256^3 f32 refcompat fused passes", flush=True)
    run_driver(tmp_path, deck)
    dg = np.loadtxt(tmp_path / "diagno.full.dat", ndmin=2)
    ds = np.loadtxt(tmp_path / "dissipation.full.dat", ndmin=2)
    lat = bgk.Lattice(n, n, n, omega=omega, storage="f32", refcompat_halos=True,
                      two_step=False)
    lat.init_hit(4.0, 0.02, 99)
    row = 0
    for it in range(1, itfin + 1):
        lat.boundaries()
        lat.col_MC()
        if it % icheck == 0:
            d = lat.diagno(refcompat=True)
            s = lat.dissipation(svisc, refcompat=True)
            assert dg[row, 0] == it and ds[row, 0] == it
            assert np.array_equal(dg[row, 1:], [d[1], d[2], d[3], d[0], d[4]]), (it, dg[row], d)
            assert ds[row, 2] == s[1] and ds[row, 3] == s[0], (it, ds[row], s)
            row += 1
    assert row == itfin // icheck
