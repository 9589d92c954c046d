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
    n1, n2, n3, n2d = task_passes(tmp_path)
    assert n1 + 2 * n2 + 3 * n3 + 2 * n2d == itfin, (n1, n2, n3, n2d)


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_driver_refcompat_fused_passes(gpu, tmp_path, storage):
    """The reference-compatible diagnostics (stale halo images) with the
    driver's fused passes (7 steps between outputs: a three-step pass and two
    two-step passes) on a 256^3 lattice: equal to the reference call order
    (boundaries + col_MC every step) exactly, for FP32 and the 16-bit kinds
    (whose three-step passes trim their second stage).  The task log's
    launch counts by pass depth show three-step passes and add up to itfin
    steps."""
    n, omega, itfin, icheck = 256, 1.95, 21, 7
    svisc = (2.0 / omega - 1.0) / 6.0
    deck = (f"&parameters\nlx={n}\nly={n}\nlz={n}\nomega={omega}\nitfin={itfin}\n"
            f"icheck={icheck}\ninit='hit'\nk0=4.0\nurms=0.02\nseed=99\n"
            f"storage='{storage}'\ndiag='refcompat'\n/\n")
    print(f"[gpu-case] fortran driver 256^3 {storage} refcompat fused passes", flush=True)
    run_driver(tmp_path, deck)
    dg = np.loadtxt(tmp_path / "diagno.full.dat", ndmin=2)
    ds = np.loadtxt(tmp_path / "dissipation.full.dat", ndmin=2)
    lat = bgk.Lattice(n, n, n, omega=omega, storage=storage, refcompat_halos=True,
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
    n1, n2, n3, n2d = task_passes(tmp_path)
    assert n3 >= itfin // icheck, (n1, n2, n3, n2d)
    assert n1 + 2 * n2 + 3 * n3 + 2 * n2d == itfin, (n1, n2, n3, n2d)


def task_passes(tmp_path):
    """Collide launches by the steps each advanced (1, 2, 3, 2 + fused
    diagnostics) from the driver's task.000000.log."""
    task = (tmp_path / "task.000000.log").read_text().splitlines()
    rows = [x for x in task if x.startswith(" # Passes 1/2/3/2+diag steps")]
    assert len(rows) == 1, task
    return [int(v) for v in rows[0][len(" # Passes 1/2/3/2+diag steps"):].split()]


def fortran_e(x, d=6, w=14):
    """Fortran Ew.d output of a double: sign, 0.<d digits>E<sign><2 digits>,
    right-justified in w characters."""
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m = f"{abs(x):.{d - 1}E}"            # d.ddddd E+xx, correctly rounded
        digits = m[0] + m[2:d + 1]
        s = ("-" if x < 0 else "") + "0." + digits + f"E{int(m.split('E')[1]) + 1:+03d}"
    return s.rjust(w)


def fixed_fields(line, widths):
    out, i = [], 0
    for wd in widths:
        out.append(line[i:i + wd])
        i += wd
    return out


def test_fortran_e_format_examples():
    assert fortran_e(0.0012345678) == "  0.123457E-02"
    assert fortran_e(-98765.4321) == " -0.987654E+05"
    assert fortran_e(1.0) == "  0.100000E+01"
    assert fortran_e(0.0) == "  0.000000E+00"
    assert fortran_e(0.9999999) == "  0.100000E+01"


@pytest.mark.gpu
def test_driver_reference_format_outputs(gpu, tmp_path):
    """The files post-processing reads, in the reference's formats:
    diagno.dat (i8, 5(e14.6,1x); SRC/diagno.F90:137,152), dissipation.dat
    (i8, 4(e14.6,1x); SRC/dissipation.F90:265-278) with the values of the
    full-precision side files rounded as Fortran's e14.6 rounds them;
    bgk.perf with two timer columns, GLUPs and the check line
    (SRC/finalize.F90:75-90); task.000000.log section timers, ratios and
    per-axis exchange lines (SRC/finalize.F90:44-63); time.log with the
    start and end dates (SRC/input.F90:39, SRC/finalize.F90:41-43)."""
    n, svisc, itfin, icheck, u0 = 24, 0.01, 60, 20, 0.02
    deck = (f"&parameters\nlx={n}\nly={n}\nlz={n}\nsvisc={svisc}\nu0={u0}\nitfin={itfin}\n"
            f"icheck={icheck}\ninit='hit'\nk0=3.0\nurms=0.02\nseed=77\n"
            f"storage='f64'\ndiag='refcompat'\n/\n")
    print("[gpu-case] fortran driver reference formats", flush=True)
    p = run_driver(tmp_path, deck)
    assert "That's all Folks" in p.stdout
    full_d = np.loadtxt(tmp_path / "diagno.full.dat", ndmin=2)
    full_s = np.loadtxt(tmp_path / "dissipation.full.dat", ndmin=2)
    dlines = (tmp_path / "diagno.dat").read_text().splitlines()
    slines = (tmp_path / "dissipation.dat").read_text().splitlines()
    assert len(dlines) == len(slines) == itfin // icheck == full_d.shape[0]
    for row, line in enumerate(dlines):
        f = fixed_fields(line, [8] + [14, 1] * 5)
        assert int(f[0]) == full_d[row, 0]
        for k in range(5):
            assert f[1 + 2 * k] == fortran_e(full_d[row, 1 + k]), (row, k, line)
            # 1x between fields (a trailing 1x ends the record unwritten)
            assert f[2 + 2 * k] == (" " if k < 4 else f[2 + 2 * k]) and f[2 + 2 * k] in (" ", "")
    for row, line in enumerate(slines):
        f = fixed_fields(line, [8] + [14, 1] * 4)
        it, t, tke, diss = full_s[row]
        assert int(f[0]) == it
        # tke / nprocs / (l m n) / u0^2 in the driver's operation order
        norm = (tke / 1.0) / float(n) / float(n) / float(n) / (u0 * u0)
        assert [f[1], f[3], f[5]] == [fortran_e(t), fortran_e(norm), fortran_e(diss)], line
    perf = (tmp_path / "bgk.perf").read_text().splitlines()
    names = ["init", "loop", "comp.", "diag.", "MPI"]
    timers = {}
    for line in perf:
        lab = line[:14].split()
        if len(lab) == 3 and lab[0] == "#" and lab[1] in names and lab[2] == "time":
            v = fixed_fields(line, [14, 14, 1, 14, 1])
            timers[lab[1]] = (float(v[1]), float(v[3]))
    assert set(timers) == set(names), perf
    glups = [line for line in perf if line.startswith(" # GLUPs")]
    check = [line for line in perf if line.startswith(" # Check")]
    assert len(glups) == 1 and len(check) == 1, perf
    loop1 = timers["loop"][1]
    if loop1 > 0:  # the GLUPs of bgk.perf use the second (time-of-day) clock
        g = float(glups[0][14:])
        assert abs(g - n ** 3 * (itfin - 1) / (loop1 * 1e9)) <= 1e-5 * g + 1e-12
    c = fixed_fields(check[0], [14, 7, 1, 7])
    assert -1.0 <= float(c[1]) <= 1.0
    task = (tmp_path / "task.000000.log").read_text()
    for key in ("# Time for section", "# Ratio", "# Ratio Coll", "# Check",
                "# Z-MPI time, BW (MB/s)", "# my up task is", "# Passes 1/2/3/2+diag steps",
                "# Pass ms 1/2/3/2+diag steps"):
        assert key in task, key
    dates = (tmp_path / "time.log").read_text().splitlines()
    assert len(dates) == 2 and all(len(x.split()) == 6 for x in dates), dates
