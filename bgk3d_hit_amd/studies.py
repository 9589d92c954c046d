"""Accuracy studies of the reduced-precision storage kinds.

    python -m bgk3d_hit_amd.studies precision --n 256 --steps 2000

runs the same seeded HIT case (k0, u_rms, omega) once per storage kind and
reports, against the FP64 run, the relative error of the kinetic-energy and
dissipation time series and of the final velocity field.

    python -m bgk3d_hit_amd.studies adjoint --n 128 --steps 500 --ckpt 25

checks the discrete-adjoint gradient of J = sum |u|^2/2 (final step) with
respect to the initial velocity against central finite differences of FP64
forward runs along seeded random solenoidal directions, and reports the
adjoint-to-forward time ratio and the adjoint's peak device memory.

Each prints one JSON document on stdout.
"""
import argparse
import json
import sys
import time

import numpy as np

from .lattice import Lattice, svisc_from_omega

C = np.array([
    (1, -1, 0), (1, 0, -1), (1, 1, 0), (1, 0, 1), (1, 0, 0), (0, 0, 1), (0, 1, 1),
    (0, 1, 0), (0, 1, -1), (-1, -1, 0), (-1, 0, -1), (-1, 1, 0), (-1, 0, 1), (-1, 0, 0),
    (0, 0, -1), (0, -1, -1), (0, -1, 0), (0, -1, 1), (0, 0, 0)], dtype=np.float64)


def velocity(f):
    """u = j / rho of shifted populations f (19, z, y, x)."""
    rho = f.sum(axis=0) + 1.0
    return np.einsum("qa,qzyx->azyx", C, f) / rho


def run_case(kind, n, steps, every, k0, urms, omega, seed, want_field=True):
    nu = svisc_from_omega(omega)
    lat = Lattice(n, n, n, omega=omega, storage=kind)
    try:
        lat.init_hit(k0, urms, seed)
        series = []
        t0 = time.perf_counter()
        for it in range(every, steps + 1, every):
            lat.step(every)
            d = lat.diagno()
            s = lat.dissipation(nu)
            series.append((it, d[0], d[4], s[0], s[1]))
        lat.synchronize()
        wall = time.perf_counter() - t0
        u = velocity(lat.get_populations()) if want_field else None
    finally:
        lat.close()
    return np.array(series), u, wall


def precision_study(n=256, steps=2000, every=100, k0=8.0, urms=0.03, omega=1.97, seed=5678,
                    kinds=("f32", "f16", "bf16"), log=None):
    ref, u_ref, wall_ref = run_case("f64", n, steps, every, k0, urms, omega, seed)
    out = {"case": {"n": n, "steps": steps, "every": every, "k0": k0, "urms": urms,
                    "omega": omega, "seed": seed},
           "f64": {"wall_s": wall_ref, "ke_final": ref[-1, 2], "diss_final": ref[-1, 3]},
           "kinds": {}}
    for kind in kinds:
        s, u, wall = run_case(kind, n, steps, every, k0, urms, omega, seed)
        ke_err = np.abs(s[:, 2] - ref[:, 2]) / np.abs(ref[:, 2])
        di_err = np.abs(s[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])
        mass_err = np.abs(s[:, 1] - ref[:, 1])
        du = u - u_ref
        field_err = float(np.sqrt((du * du).sum() / (u_ref * u_ref).sum()))
        out["kinds"][kind] = {
            "ke_rel_err_max": float(ke_err.max()), "ke_rel_err_final": float(ke_err[-1]),
            "diss_rel_err_max": float(di_err.max()), "diss_rel_err_final": float(di_err[-1]),
            "mean_rho_abs_err_max": float(mass_err.max()),
            "velocity_field_rel_l2_err_final": field_err,
            "wall_s": wall}
        if log:
            log(f"[precision] {kind}: KE {ke_err.max():.2e} diss {di_err.max():.2e} "
                f"field {field_err:.2e}")
    return out


def _timed_forward(lat, u0, steps):
    """Plain forward: device time of `steps` fused steps plus the final
    kinetic-energy reduction (initialisation excluded)."""
    lat.init_equilibrium(u0)
    lat.synchronize()
    t0 = time.perf_counter()
    lat.step(steps)
    ke = lat.diagno()[4]
    return (time.perf_counter() - t0) * 1e3, ke


def adjoint_study(n=128, steps=500, ckpt=25, levels=1, k0=4.0, urms=0.02, omega=1.90, seed=42,
                  ndir=16, eps=1e-3, kind="f32", fd=True, log=None):
    from .lattice import hit_velocity
    u0 = hit_velocity(n, n, n, k0, urms, seed)
    lat = Lattice(n, n, n, omega=omega, storage=kind)
    try:
        _timed_forward(lat, u0, min(steps, 20))  # warm-up
        fwd_ms, _ = _timed_forward(lat, u0, steps)
        t0 = time.perf_counter()
        J, g, st = lat.adjoint_ke(u0, steps, ckpt_every=ckpt, levels=levels)
        adj_wall_ms = (time.perf_counter() - t0) * 1e3
        lat_bytes = lat.info()["device_bytes"]
    finally:
        lat.close()
    out = {"case": {"n": n, "steps": steps, "ckpt_every": ckpt, "levels": levels, "k0": k0,
                    "urms": urms, "omega": omega, "seed": seed, "kind": kind},
           "J": J, "forward_ms": fwd_ms,
           "adjoint_forward_ms": st["forward_ms"], "adjoint_backward_ms": st["backward_ms"],
           "adjoint_total_ms": st["forward_ms"] + st["backward_ms"],
           "adjoint_call_wall_ms": adj_wall_ms,
           "ratio_adjoint_to_forward": (st["forward_ms"] + st["backward_ms"]) / fwd_ms,
           "recomputed_steps": st["recomputed_steps"],
           "adjoint_peak_bytes": st["peak_bytes"],
           "lattice_bytes": lat_bytes,
           "peak_hbm_bytes": st["peak_bytes"] + lat_bytes}
    # kernel time of the taped forward passes (re-taping overlaps the
    # adjoint steps on a second stream) and of the adjoint-step passes, from
    # a second run with HIP events around every launch
    lat = Lattice(n, n, n, omega=omega, storage=kind, timing=True)
    try:
        lat.adjoint_ke(u0, steps, ckpt_every=ckpt, levels=levels)
        tk = lat.timing_adjoint()
    finally:
        lat.close()
    out["taped_forward_kernel_ms"] = tk["tape"][0]
    out["taped_forward_launches"] = tk["tape"][1]
    out["adjoint_step_kernel_ms"] = tk["adjoint"][0]
    out["adjoint_step_launches"] = tk["adjoint"][1]
    if log:
        log(f"[adjoint] {n}^3 x {steps} ckpt {ckpt} levels {levels}: ratio "
            f"{out['ratio_adjoint_to_forward']:.2f}, peak {out['peak_hbm_bytes'] / 2**30:.1f} GiB")
    if fd:
        ref = Lattice(n, n, n, omega=omega, storage="f64")
        errs = []
        try:
            for i in range(ndir):
                d = hit_velocity(n, n, n, k0, urms, 1000 + i)
                jp = ref.forward_ke(u0 + eps * d, steps)
                jm = ref.forward_ke(u0 - eps * d, steps)
                fdv = (jp - jm) / (2 * eps)
                ad = float((g * d).sum())
                errs.append(abs(ad - fdv) / abs(fdv))
        finally:
            ref.close()
        out["fd"] = {"directions": ndir, "eps": eps, "rel_err": errs,
                     "rel_err_max": max(errs)}
        if log:
            log(f"[adjoint] FD along {ndir} directions: max rel err {max(errs):.2e}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("precision")
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--every", type=int, default=100)
    p.add_argument("--k0", type=float, default=8.0)
    p.add_argument("--urms", type=float, default=0.03)
    p.add_argument("--omega", type=float, default=1.97)
    p.add_argument("--seed", type=int, default=5678)
    p.add_argument("--kinds", default="f32,f16,bf16")
    q = sub.add_parser("adjoint")
    q.add_argument("--n", type=int, default=128)
    q.add_argument("--steps", type=int, default=500)
    q.add_argument("--ckpt", type=int, default=25)
    q.add_argument("--levels", type=int, default=1)
    q.add_argument("--k0", type=float, default=4.0)
    q.add_argument("--urms", type=float, default=0.02)
    q.add_argument("--omega", type=float, default=1.90)
    q.add_argument("--seed", type=int, default=42)
    q.add_argument("--ndir", type=int, default=16)
    q.add_argument("--eps", type=float, default=1e-3)
    q.add_argument("--kind", default="f32")
    q.add_argument("--no-fd", action="store_true")
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "precision":
        res = precision_study(a.n, a.steps, a.every, a.k0, a.urms, a.omega, a.seed,
                              tuple(k for k in a.kinds.split(",") if k), log=log)
    else:
        res = adjoint_study(a.n, a.steps, a.ckpt, a.levels, a.k0, a.urms, a.omega, a.seed,
                            a.ndir, a.eps, a.kind, fd=not a.no_fd, log=log)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
