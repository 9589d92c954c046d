"""Discrete adjoint on the GPU: against the oracle's FP64 adjoint (a
generic-table restatement, itself pinned by finite differences on the CPU),
against central finite differences of the GPU forward, across checkpoint
levels, and through the torch.autograd binding."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import bgk3d_hit_amd as bgk
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.mark.parametrize("levels,K", [(0, 1), (1, 5), (2, 3)])
def test_adjoint_f64_matches_oracle(gpu, levels, K):
    shape, omega, T = (12, 10, 8), 1.8, 13
    say("adjoint vs oracle", levels, K)
    u0 = bgk.hit_velocity(*shape, 2.0, 0.03, 21)
    J_o, g_o = oracle.adjoint_ke(*shape, omega, T, u0)
    lat = bgk.Lattice(*shape, omega=omega, storage="f64")
    J, g, st = lat.adjoint_ke(u0, T, ckpt_every=K, levels=levels)
    assert abs(J - J_o) < 1e-13 * J_o
    assert np.abs(g - g_o).max() < 1e-10 * np.abs(g_o).max()
    if levels == 0:
        assert st["recomputed_steps"] == 0


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix"])
@pytest.mark.parametrize("levels,K", [(0, 1), (1, 7), (2, 3), (1, 5)])
@pytest.mark.parametrize("three", [True, False])
def test_adjoint_two_step_passes_equal_single_steps(gpu, kind, levels, K, three):
    """Taped forward and adjoint sweep with three and two steps per pass
    (forced at a small size; segment lengths 1, 3, 5, 7 and a 31-step full
    tape, so fused passes of both depths and single steps mix) give the
    single-step gradient bit for bit."""
    shape, omega, T = (70, 9, 20), 1.9, 31
    say("adjoint fused vs single", kind, levels, K, three)
    u0 = bgk.hit_velocity(*shape, 2.0, 0.03, 8)
    a = bgk.Lattice(*shape, omega=omega, storage=kind, two_step="always", three_step=three)
    b = bgk.Lattice(*shape, omega=omega, storage=kind, two_step=False)
    Ja, ga, _ = a.adjoint_ke(u0, T, ckpt_every=K, levels=levels)
    Jb, gb, _ = b.adjoint_ke(u0, T, ckpt_every=K, levels=levels)
    assert Ja == Jb
    assert np.array_equal(ga, gb)


@pytest.mark.parametrize("kind", ["f64", "f32", "f32mix"])
def test_adjoint_wide_offsets_equal(gpu, kind):
    """The 64-bit-offset kernels of the adjoint (taped forward, two-step
    adjoint sweep, forced by BGK_FLAG_WIDE_OFFSETS) give the same gradient."""
    shape, omega, T = (70, 9, 20), 1.9, 13
    say("adjoint wide offsets", kind)
    u0 = bgk.hit_velocity(*shape, 2.0, 0.03, 8)
    a = bgk.Lattice(*shape, omega=omega, storage=kind, two_step="always")
    b = bgk.Lattice(*shape, omega=omega, storage=kind, two_step="always", wide_offsets=True)
    Ja, ga, _ = a.adjoint_ke(u0, T, ckpt_every=5, levels=1)
    Jb, gb, _ = b.adjoint_ke(u0, T, ckpt_every=5, levels=1)
    assert Ja == Jb
    assert np.array_equal(ga, gb)


def test_adjoint_checkpoint_levels_agree_bitwise(gpu):
    shape, omega, T = (24, 24, 24), 1.9, 30
    say("adjoint levels bitwise")
    u0 = bgk.hit_velocity(*shape, 3.0, 0.02, 5)
    lat = bgk.Lattice(*shape, omega=omega, storage="f32")
    res = [lat.adjoint_ke(u0, T, ckpt_every=K, levels=lv) for lv, K in ((0, 1), (1, 4), (2, 4))]
    for J, g, st in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(g, res[0][1])
    # one level: every segment but the last two (taped by the forward) is
    # re-run: segments [0,4) .. [20,24) of [0,30)
    assert res[1][2]["recomputed_steps"] == ((T + 3) // 4 - 2) * 4
    # two levels: 3 coarse segments of 12 steps; the same 24 re-taped steps
    # and the fine checkpoints of the first two coarse segments re-run
    # (8 steps each)
    assert res[2][2]["recomputed_steps"] == 40


def test_adjoint_central_differences_fp64(gpu):
    shape, omega, T = (32, 32, 32), 1.9, 100
    say("adjoint FD fp64")
    u0 = bgk.hit_velocity(*shape, 4.0, 0.02, 42)
    lat = bgk.Lattice(*shape, omega=omega, storage="f64")
    J, g, _ = lat.adjoint_ke(u0, T, ckpt_every=25, levels=1)
    assert abs(lat.forward_ke(u0, T) - J) < 1e-13 * J
    eps = 1e-3
    for s in range(1, 5):
        d = bgk.hit_velocity(*shape, 4.0, 0.02, 1000 + s)
        fd = (lat.forward_ke(u0 + eps * d, T) - lat.forward_ke(u0 - eps * d, T)) / (2 * eps)
        ad = float((g * d).sum())
        assert abs(fd - ad) < 1e-6 * abs(fd), (s, fd, ad)


def test_adjoint_fp32_against_fp64_cpu_adjoint(gpu):
    """FP32 GPU adjoint against the FP64 CPU adjoint at 32^3."""
    shape, omega, T = (32, 32, 32), 1.90, 60
    say("adjoint fp32 vs cpu fp64")
    u0 = bgk.hit_velocity(*shape, 4.0, 0.02, 42)
    J_o, g_o = oracle.adjoint_ke(*shape, omega, T, u0)
    lat = bgk.Lattice(*shape, omega=omega, storage="f32")
    J, g, _ = lat.adjoint_ke(u0, T, ckpt_every=25, levels=1)
    assert abs(J - J_o) < 1e-5 * J_o
    err = np.sqrt(((g - g_o) ** 2).sum() / (g_o ** 2).sum())
    assert err < 1e-4, err


def test_torch_autograd_binding(gpu):
    """torch.autograd.Function over the device-pointer adjoint.  Runs in a
    child process that imports torch before the library (one HIP runtime)."""
    say("torch autograd (child process)")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torch_adjoint_check.py")
    p = subprocess.run([sys.executable, "-u", script], capture_output=True, text=True,
                       timeout=280)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert res["J_rel_err"] < 1e-14
    assert res["grad_equal"]


def test_adjoint_128cubed_500_steps_against_finite_differences(gpu):
    """128^3 FP32, k0=4, u_rms=0.02, omega=1.90, 500 steps, checkpoints
    every 25: gradient within 1e-4 of FP64 central differences (eps=1e-3)
    along 16 seeded solenoidal directions."""
    from bgk3d_hit_amd.studies import adjoint_study
    say("adjoint 128^3 x 500 FD x16")
    res = adjoint_study(n=128, steps=500, ckpt=25, levels=1, ndir=16, eps=1e-3, kind="f32")
    print(res["ratio_adjoint_to_forward"], res["fd"]["rel_err_max"], flush=True)
    assert res["fd"]["rel_err_max"] < 1e-4
