"""Benchmark of the BGK3D_HIT hot path on one MI355X.

Workload (one "step" = one stream-collide time step of the lattice):
512^3 triply periodic D3Q19 BGK, seeded HIT initial field (k0 = 4,
u_rms = 0.02, seed 1234), omega = 1.95, with the energy/dissipation
reductions (diagno + dissipation) every 50 steps inside the timed region.
The headline `value` is the FP32 run (storage and compute FP32); the FP64
and reduced-storage (FP16 / BF16 shifted f_i - w_i, FP32 collide) runs of
the same workload are reported under `variants`.

Each block of steps and its reductions is one bgk_step_diagnostics call:
lattices this size advance three time steps per launch (k_collide3: the
first two steps' states kept in LDS), and the last two steps of a block
(50 = 16 * 3 + 2) share one pass with the diagnostics of the state they
leave (k_collide3's RED form: two steps, then diagno and the pulled
dissipation sums in place of a third collision) instead of a two-step pass
followed by a reduction pass.  The two-step pass alone and the
one-pass-per-step kernel (k_collide) are reported beside it (their blocks
end in a separate reduction pass).

`roofline` is measured live for the dominant kernel (the launch depth with
the most collide time): HIP events around every collide launch on the stream
it runs on give the average launch duration.  `achieved` counts the bytes
one launch must move: every population of every site once in and once out,
19 loads + 19 stores = 38 * s bytes per site, 38 * s * 512^3 per launch
whether the launch advances one, two or three time steps (the fused passes
keep their intermediate states on chip), so `frac` = achieved / 8 TB/s is
at most 1.  `site_update_equivalent_gbs` counts the same launch per
site-update (38 * s per site and time step, times the steps it advances):
the rate a one-step-per-launch kernel would need to match it, which exceeds
the HBM peak for the fused passes.  The reductions run every 50 steps and
after the last timed step, so any --steps times at least one of them
(`config.reductions_in_timed_region`).  `traffic` is the HBM bytes per launch measured by rocprofv3
FETCH_SIZE / WRITE_SIZE runs of this script (profiles/collide_traffic.json,
null when absent or for another lattice size); `hbm_actual_gbs` = traffic /
launch time is the kernel's own DRAM rate, and `copy_ceiling_gbs` the rate
of a plain copy of the same 38-stream pattern in the same layout
(bin/stream_ceiling, profiles/stream_ceiling_512cubed.jsonl).

`cpu_baseline` times the CPU oracle (plain C/OpenMP restatement of the
reference Fortran) on a bounded sample on the host cores.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + % HBM roofline, 512^3 D3Q19 (FP64/FP32/FP16-store); adjoint/fwd time"
HBM_PEAK_GBS = 8000.0
BYTES = {"f64": 8, "f32": 4, "f32mix": 4, "f16": 2, "bf16": 2}
COMPUTE = {"f64": "f64", "f32": "f32", "f32mix": "f64", "f16": "f32", "bf16": "f32"}
KERNEL = {1: "k_collide", 2: "k_collide2", 3: "k_collide3"}  # by time steps per launch


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(kind, n, steps, warmup, diag_every, omega=1.95, shape=None, layout="periodic",
            k0=4.0, urms=0.02, seed=1234, two_step=True, three_step=True):
    import bgk3d_hit_amd as bgk

    shape = shape or (n, n, n)
    nu = bgk.svisc_from_omega(omega)
    lat = bgk.Lattice(*shape, omega=omega, storage=kind, layout=layout, timing=True,
                      two_step=two_step, three_step=three_step)
    try:
        lat.init_hit(k0, urms, seed)
        d0 = lat.diagno()
        lat.step(warmup)
        lat.synchronize()
        lat.timing(reset=True)
        diag = []
        t0 = time.perf_counter()
        done = 0
        # the reductions run every diag_every steps and after the last
        # step, so any step count times at least one of them:
        # ceil(steps / diag_every) diagnostics passes in the timed region
        while done < steps:
            k = min(diag_every, steps - done)
            # k steps, then diagno + dissipation of the state they leave
            d5, s2 = lat.step_diagnostics(k, nu)
            done += k
            diag.append((done + warmup, *d5, *s2))
        lat.synchronize()
        wall = time.perf_counter() - t0
        passes = lat.timing_passes(reset=False)
        fused = lat.timing_fused(reset=False)
        t = lat.timing(reset=True)
    finally:
        lat.close()
    nsite = shape[0] * shape[1] * shape[2]
    launches = max(t["collide_launches"], 1)
    per_step = launches / steps
    step_collide_ms = t["collide_ms"] / steps
    # the dominant kernel: the launch depth (time steps per launch) with the
    # most collide time in the timed region
    spl = max(passes, key=lambda k: passes[k][0])
    avg_ms = passes[spl][0] / max(passes[spl][1], 1)
    # algorithmic bytes one launch must move: every population of every site
    # once in and once out, 38 * s per site, whatever the number of time
    # steps the launch advances (the fused passes keep their intermediate
    # states on chip)
    algo_bytes = 38 * BYTES[kind] * nsite
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    # the same launch counted per site-update (38 * s per site and time
    # step): the rate a one-step-per-launch kernel would need to match it
    su_bytes = algo_bytes * spl
    su_gbs = su_bytes / (avg_ms * 1e-3) / 1e9
    return {
        "kind": kind,
        "kernel": KERNEL[spl] + f"<{kind}>",
        "steps_per_launch": spl,
        "launches_by_depth": {str(k): v[1] for k, v in passes.items() if v[1]},
        "mlups": nsite * steps / wall / 1e6,
        "ms_per_step": wall * 1e3 / steps,
        "collide_avg_ms": avg_ms,
        "collide_ms_per_step": step_collide_ms,
        "collide_launches_per_step": per_step,
        "collide_launches": t["collide_launches"],
        "halo_ms_per_step": t["halo_ms"] / steps,
        "reduce_ms_total": t["reduce_ms"],
        "fused_pass_avg_ms": fused[0] / max(fused[1], 1),
        "fused_passes": fused[1],
        "achieved_gbs": achieved,
        "frac": achieved / HBM_PEAK_GBS,
        "algo_bytes_per_launch": algo_bytes,
        "site_update_bytes_per_launch": su_bytes,
        "site_update_equivalent_gbs": su_gbs,
        "reductions_timed": len(diag),
        "nsite": nsite,
        "ke0": d0[4],
        "diag_last": diag[-1] if diag else None,
    }


def summary(r, ref=None):
    out = {"mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms_per_step"], 4),
           "collide_ms_per_step": round(r["collide_ms_per_step"], 4),
           "kernel": r["kernel"], "steps_per_launch": r["steps_per_launch"],
           "achieved_gbs": round(r["achieved_gbs"], 1), "frac": round(r["frac"], 4),
           "site_update_equivalent_gbs": round(r["site_update_equivalent_gbs"], 1),
           "storage_bytes": BYTES[r["kind"]], "compute": COMPUTE[r["kind"]],
           "traffic": load_traffic(r["kind"], r["steps_per_launch"], r["nsite"])}
    if out["traffic"]:
        out["hbm_actual_gbs"] = round(out["traffic"] / (r["collide_avg_ms"] * 1e-3) / 1e9, 1)
    if ref is not None:
        out["speedup_vs_f32"] = round(r["mlups"] / ref["mlups"], 3)
    return out


def guarded(name, fn):
    try:
        return fn()
    except Exception as e:  # an extra configuration must not cost the headline line
        log(f"[bench] {name} failed: {e}")
        return {"error": str(e)[:300]}


def slab_config(steps, warmup):
    out = {"workload": "one rank's slab 1024x1024x256 FP32, +-z ghost planes, pack / "
                       "self-exchange / unpack of the 5 crossing populations per face, "
                       "bulk || boundary on two HIP streams, k0=8 u_rms=0.03 omega=1.97"}
    paths = {
        "three_step_pass": ({}, "three-step pass on planes 3..252 (stream 0) || halo exchange, "
                                "boundary planes stepped three times through two plane buffers "
                                "(stream 1); 16 passes and a two-step pass per 50 steps"),
        "two_step_pass": ({"three_step": False},
                          "two-step pass on planes 2..253 (stream 0) || halo exchange, "
                          "boundary planes stepped twice through a plane buffer (stream 1)"),
        "one_pass_per_step": ({"two_step": False},
                              "one-step kernel on planes 1..254 (stream 0) || halo exchange "
                              "and boundary planes (stream 1)"),
    }
    for name, (kw, path) in paths.items():
        r = run_gpu("f32", 0, steps, warmup, 50, omega=1.97, shape=(1024, 1024, 256),
                    layout="slab", k0=8.0, urms=0.03, seed=5678, **kw)
        nsite = 1024 * 1024 * 256
        # two streams overlap, so the collide time summed over both is not a
        # kernel rate: report the whole step's one-step-equivalent rate
        out[name] = {
            "mlups": round(r["mlups"], 1), "ms_per_step": round(r["ms_per_step"], 4),
            "path": path,
            "one_step_equivalent_gbs": round(38 * 4 * nsite / (r["ms_per_step"] * 1e-3) / 1e9, 1),
            "halo_ms_per_step": round(r["halo_ms_per_step"], 4)}
        log(f"[bench] slab {name}: {r['mlups']:.0f} MLUPS")
    return out


def periodic_1024_config(steps, warmup):
    res = {}
    ref = None
    for kind in ("f32", "f16", "bf16"):
        r = run_gpu(kind, 1024, steps, warmup, 50, omega=1.97, k0=8.0, urms=0.03, seed=5678)
        if kind == "f32":
            ref = r
        res[kind] = summary(r, ref)
        log(f"[bench] 1024^3 {kind}: {r['mlups']:.0f} MLUPS")
    res["workload"] = "1024^3 periodic HIT k0=8 u_rms=0.03 omega=1.97, shifted storage"
    return res


def adjoint_128_config():
    from bgk3d_hit_amd.studies import adjoint_study
    r = adjoint_study(n=128, steps=500, ckpt=25, levels=1, fd=False)
    return {"workload": "128^3 FP32 HIT k0=4 u_rms=0.02 omega=1.90, 500-step horizon, "
                        "checkpoints every 25 steps with recompute; the backward re-tapes "
                        "the next segment on a second stream beside the adjoint steps",
            "forward_ms": round(r["forward_ms"], 2),
            "adjoint_forward_ms": round(r["adjoint_forward_ms"], 2),
            "adjoint_backward_ms": round(r["adjoint_backward_ms"], 2),
            "retape_and_last_segment_tape_kernel_ms": round(r["taped_forward_kernel_ms"], 2),
            "adjoint_step_kernel_ms": round(r["adjoint_step_kernel_ms"], 2),
            "adjoint_total_ms": round(r["adjoint_total_ms"], 2),
            "ratio_adjoint_to_forward": round(r["ratio_adjoint_to_forward"], 3),
            "peak_hbm_bytes": r["peak_hbm_bytes"]}


def cpu_baseline(min_seconds=10.0, max_steps=400):
    import numpy as np

    import bgk3d_hit_amd as bgk
    import oracle

    n = 256
    u = bgk.hit_velocity(n, n, n, 4.0, 0.02, 1234)
    o = oracle.OracleLattice(n, n, n, omega=1.95, kind="f32")
    o.init_equilibrium(u)
    del u
    o.step(1)  # first touch of every page
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps and (time.perf_counter() - t0) < min_seconds:
        o.step(1)
        steps += 1
    wall = time.perf_counter() - t0
    o.close()
    return {"value": round(n ** 3 * steps / wall / 1e6, 2), "unit": "MLUPS",
            "cores": oracle.num_threads(), "kind": "port",
            "sample": f"CPU oracle (C/OpenMP restatement of SRC/col_MC.F90 + "
                      f"SRC/bcond.comm.step3.F90), FP32, 256^3 HIT sample of the same "
                      f"workload, {steps} steps in {wall:.1f} s"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_samples(seconds=2.0):
    """The oracle at the small configurations (64^3 FP64 as in the CPU-runnable
    case, 128^3 FP32 and FP64), on all of its threads and on one thread, each
    for about `seconds`; MLUPS per sample."""
    import bgk3d_hit_amd as bgk
    import oracle

    nthreads = oracle.num_threads()
    out = {"cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "samples": []}
    for n, kind in ((64, "f64"), (128, "f32"), (128, "f64")):
        u = bgk.hit_velocity(n, n, n, 4.0, 0.02, 1234)
        for threads in (nthreads, 1):
            oracle.set_num_threads(threads)
            o = oracle.OracleLattice(n, n, n, omega=1.90, kind=kind)
            o.init_equilibrium(u)
            o.step(1)
            steps, t0 = 0, time.perf_counter()
            while steps < 2 or (time.perf_counter() - t0) < seconds:
                o.step(1)
                steps += 1
            wall = time.perf_counter() - t0
            o.close()
            out["samples"].append({"lattice": n, "kind": kind, "threads": threads,
                                   "steps": steps,
                                   "mlups": round(n ** 3 * steps / wall / 1e6, 2)})
    oracle.set_num_threads(nthreads)
    return out


def calibrate_copy(gib, reps=10):
    """Plain copy of a known size: dst.copy_(src) moves gib GiB in and gib GiB out."""
    import torch

    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 2 * n * 4
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4,
                      "avg_ms": ms, "gbs": nbytes / (ms * 1e-3) / 1e9}), flush=True)


def load_traffic(kind, steps_per_launch, nsite):
    """HBM bytes per launch of the collide kernel at 512^3 (rocprofv3 counters)."""
    if nsite != 512 ** 3:
        return None
    p = os.path.join(ROOT, "profiles", "collide_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        key = KERNEL[steps_per_launch] + "/" + kind
        return d.get(key, {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def copy_ceiling(nsite):
    """GB/s of a plain 19-in/19-out copy in the interleaved layout at 512^3
    (bin/stream_ceiling), the achievable rate for the 38-stream pattern."""
    if nsite != 512 ** 3:
        return None
    p = os.path.join(ROOT, "profiles", "stream_ceiling_512cubed.jsonl")
    try:
        with open(p) as f:
            for line in f:
                d = json.loads(line)
                if d.get("kernel") == "copy19_interleaved":
                    return d.get("gbs")
    except (OSError, ValueError):
        pass
    return None


def roofline(r):
    traffic = load_traffic(r["kind"], r["steps_per_launch"], r["nsite"])
    out = {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1),
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(r["frac"], 4),
           "traffic": traffic, "kernel": r["kernel"],
           "steps_per_launch": r["steps_per_launch"],
           "avg_launch_ms": round(r["collide_avg_ms"], 4),
           "algo_bytes_per_launch": r["algo_bytes_per_launch"],
           "site_update_bytes_per_launch": r["site_update_bytes_per_launch"],
           "site_update_equivalent_gbs": round(r["site_update_equivalent_gbs"], 1),
           "site_update_equivalent_frac": round(r["site_update_equivalent_gbs"] / HBM_PEAK_GBS,
                                                4),
           "hbm_actual_gbs": None, "hbm_actual_frac": None}
    if traffic:
        act = traffic / (r["collide_avg_ms"] * 1e-3) / 1e9
        out["hbm_actual_gbs"] = round(act, 1)
        out["hbm_actual_frac"] = round(act / HBM_PEAK_GBS, 4)
    ceil = copy_ceiling(r["nsite"])
    out["copy_ceiling_gbs"] = ceil
    out["hbm_actual_frac_of_copy"] = (round(out["hbm_actual_gbs"] / ceil, 4)
                                      if ceil and out["hbm_actual_gbs"] else None)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--diag-every", type=int, default=50)
    ap.add_argument("--variants", default="f64,f16,bf16",
                    help="extra storage kinds to measure beside f32 ('' for none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the slab, 1024^3 and adjoint configurations")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="GIB",
                    help="only time a plain device copy of GIB GiB (the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    a = ap.parse_args()
    if a.calibrate_copy > 0:
        calibrate_copy(a.calibrate_copy)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the one-GPU throughput is the whole-job value", file=sys.stderr)
        sys.exit(2)
    import bgk3d_hit_amd as bgk

    if bgk.device_count() < 1:
        log("bench.py: no GPU visible")
        sys.exit(1)
    log(f"[bench] f32 {a.size}^3 warmup {a.warmup} steps {a.steps}")
    main_run = run_gpu("f32", a.size, a.steps, a.warmup, a.diag_every)
    log(f"[bench] f32: {main_run['mlups']:.0f} MLUPS, {main_run['kernel']} "
        f"{main_run['collide_avg_ms']:.3f} ms/launch = {main_run['achieved_gbs']:.0f} GB/s")
    variants = {}
    refs = {}  # the f32 run of each mode, the reference of the other kinds' speedups
    modes = (("", {}), ("_two_step_pass", {"three_step": False}),
             ("_one_pass_per_step", {"two_step": False}))
    for kind in ["f32"] + [k for k in a.variants.split(",") if k and k != "f32"]:
        depth = 3
        for suffix, kw in modes:
            name = kind + suffix
            if kind == "f32" and not suffix:
                refs[suffix] = main_run
                depth = main_run["steps_per_launch"]
                continue
            if (suffix == "_two_step_pass" and depth < 3) or (suffix and depth < 2):
                continue  # the same kernel as the default run
            log(f"[bench] {name} {a.size}^3")
            r = guarded(name, lambda: run_gpu(kind, a.size, a.steps, a.warmup, a.diag_every,
                                              **kw))
            if "error" in r:
                variants[name] = r
                continue
            log(f"[bench] {name}: {r['mlups']:.0f} MLUPS, {r['kernel']} {r['achieved_gbs']:.0f} GB/s")
            if kind == "f32":
                refs[suffix] = r
            if not suffix:
                depth = r["steps_per_launch"]
            variants[name] = summary(r, refs.get(suffix))
    configs = {}
    if not a.no_extra:
        log("[bench] slab 1024x1024x256")
        configs["slab_1024x1024x256"] = guarded("slab", lambda: slab_config(100, 10))
        log("[bench] 1024^3 f32/f16/bf16")
        configs["periodic_1024cubed"] = guarded("1024^3", lambda: periodic_1024_config(100, 10))
        log("[bench] adjoint 128^3")
        configs["adjoint_128cubed"] = guarded("adjoint", adjoint_128_config)
    cpu = None
    cpu_more = None
    if not a.no_cpu:
        log("[bench] cpu baseline")
        cpu = guarded("cpu", cpu_baseline)
        log(f"[bench] cpu: {cpu}")
        cpu_more = guarded("cpu samples", cpu_samples)
        log(f"[bench] cpu samples: {cpu_more}")
    out = {
        "metric": METRIC,
        "value": round(main_run["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded HIT velocity (k0=4, u_rms=0.02, seed 1234), rho=1 "
                "equilibrium populations",
        "config": {"workload": f"{a.size}^3 periodic D3Q19 BGK HIT, omega=1.95, FP32 storage "
                               f"and collide, diagno+dissipation every {a.diag_every} steps "
                               f"and after the last step",
                   "lattice": [a.size] * 3, "omega": 1.95, "diag_every": a.diag_every,
                   "reductions_in_timed_region": main_run["reductions_timed"],
                   "reductions_fused_with_last_two_steps": main_run["fused_passes"],
                   "fused_pass_avg_ms": round(main_run["fused_pass_avg_ms"], 4),
                   "layout": "per copy [z][y][q][x] (19 population rows of a y line side by "
                             "side), halo-free, periodic wrap in index math"},
        "roofline": roofline(main_run),
        "cpu_baseline": cpu,
        "cpu_samples": cpu_more,
        "variants": variants,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
