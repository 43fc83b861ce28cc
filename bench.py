"""Benchmark of the Kobayashi step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--kernel fused|two_pass]

Workload (the single-large-grid configuration): 4096 x 4096 FP32 grid,
zero-flux boundaries, dx = 0.03, dt = 1e-4, tau = 3e-4, eps_bar = 0.01,
delta = 0.04, j = 6, theta0 = 0.2, alpha = 0.9, gamma = 10, K = 2.0, Teq = 1,
noise a = 0.01, nine r = 5 nuclei at Philox-drawn centres (seed 0, margin 8).
A step is one explicit-Euler step of the whole grid; the timed region is K
steps of that trajectory starting from its initial state (default K = 20000,
the configuration's full horizon), with the fields resident in HBM.

Prints one JSON line.  `value` is cell-updates per second of the timed
region; `roofline` prices the step's average duration on the GPU (HIP events
on the solver's own stream around the timed steps: each step is one launch
of the fused kernel, over every 2 x 4-tile chunk or over the live ones; every
7th launch also answers the host's liveness request) at 16 algorithmic bytes per
cell-update against the 8 TB/s HBM peak; `cpu_baseline` times the CPU oracle (the reference's
arithmetic restated in C, one thread) on a bounded sample of the same
workload (`cpu_all_cores`: the same oracle on OpenMP threads, labelled
apart).  Beside the headline it reports:
  batched   -- the 512-member ensemble (sides 128..1024, per-member
               parameters) over its 5000-step horizon from disk nuclei: one
               batched launch per step (autotuned launch mode) vs looping the
               plain solver per member (one handle and stream per member)
  adjoint   -- loss + gradients of 1000 steps on 512^2 (checkpoint every 50)
               vs the plain forward of the same steps
  dense_kernel -- the same 4096^2 trajectory on the fused launch without tile
               status (KG_KERNEL_FUSED_DENSE, bitwise the same results)
  reference_config -- the reference's configuration (300^2 FP64, 2000 steps)
               on the GPU and on the CPU oracle, compared cell by cell
--calibrate-copy runs only a device copy of known size (for scaling the
rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s & %HBM roofline; batched/adjoint vs plain at stated FP64 err"
HBM_PEAK_GBS = 8000.0
ALG_BYTES_FP32 = 16  # phi and T read once and written once per cell-update
WORKLOAD = dict(delta=0.04, aniso=6.0, K=2.0, theta0=0.2, noise_a=0.01, boundary=1,
                seed=0)            # the 4096^2 configuration (crystalgrowth_amd/configs.py c1)
SEEDS = dict(count=9, r=5.0, margin=8, seed=0)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--kernel", default="fused", choices=["fused", "fused_dense", "two_pass"])
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="steps of the CPU oracle sample (0 skips the CPU baseline)")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=5000,
                    help="steps of the batched-vs-looped comparison (0 skips it)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-looped", action="store_true",
                    help="batched leg without the looped plain handles (for profiling)")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the same trajectory on the fused kernel without tile status")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--sync-every", type=int, default=0,
                    help="synchronise after every N timed steps (for counter-collection "
                         "profiler runs; 0 = queue all steps)")
    return ap.parse_args()


def torch_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def traffic_from_profiles(size: int):
    """HBM bytes per launch of the fused step measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and
    recorded under profiles/; None when no measurement exists."""
    path = os.path.join(ROOT, "profiles", f"kobayashi_step_{size}x{size}_fp32_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(size: int, nsteps: int):
    from oracle import oracle as O
    p = O.default_params(**WORKLOAD)
    s, _ = O.init_seeds(size, size, SEEDS["count"], SEEDS["r"], SEEDS["margin"], SEEDS["seed"],
                        dtype=np.float32)
    t0 = time.perf_counter()
    O.run(s, p, nsteps)
    dt = time.perf_counter() - t0
    cups = size * size * nsteps / dt
    return {"value": cups / 1e9, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle (reference float arithmetic restated in C, gcc -O2, 1 thread) "
                      f"on the {size}x{size} workload, steps 0-{nsteps - 1} from the seeded "
                      f"initial state, {dt:.1f} s on '{cpu_model()}' ({os.cpu_count()} cores)"}


def cpu_all_cores(size: int, nsteps: int):
    """The same oracle with its row loops on OpenMP threads (bitwise the
    single-thread results; tests/test_oracle_parallel.py), labelled apart
    from cpu_baseline: the reference itself runs on one thread."""
    from oracle import oracle as O
    p = O.default_params(**WORKLOAD)
    s, _ = O.init_seeds(size, size, SEEDS["count"], SEEDS["r"], SEEDS["margin"], SEEDS["seed"],
                        dtype=np.float32)
    threads = O.threads_parallel()
    t0 = time.perf_counter()
    O.run_parallel(s, p, nsteps)
    dt = time.perf_counter() - t0
    return {"value": size * size * nsteps / dt / 1e9, "unit": "Gcell-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"CPU oracle built with OpenMP over rows, {threads} threads, on the "
                      f"{size}x{size} workload, steps 0-{nsteps - 1}, {dt:.1f} s"}


def reference_config(kb, cpu_steps: int):
    """The reference's own configuration (300^2 FP64, disk r = 5, 2000 steps,
    crystalgrowth_amd/configs.py c0): the GPU run, the CPU oracle's full run
    (OpenMP build, bitwise the serial oracle) compared per cell, and the
    oracle's single-thread rate on its first cpu_steps steps."""
    from crystalgrowth_amd import configs
    from oracle import oracle as O
    c = configs.C0
    n, steps = c["n"], c["steps"]
    p = configs.c0_params()
    with kb.Solver(n, n, "f64", p) as s:
        s.reset_disk(n // 2, n // 2, c["disk_r"])
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        s.synchronize()
        tg = time.perf_counter() - t0
        gphi, gT = s.phase(), s.temperature()
    op = O.default_params(**{k: getattr(p, k) for k in ("dx", "dy", "dt", "tau", "eps_bar", "K",
                                                         "delta", "aniso", "alpha", "gamma",
                                                         "teq", "theta0", "noise_a",
                                                         "boundary", "seed")})
    st = O.init_disk(n, n, n // 2, n // 2, c["disk_r"], np.float64)
    t0 = time.perf_counter()
    ref = O.run_parallel(st, op, steps)
    tp = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.run(st, op, cpu_steps)
    t1 = time.perf_counter() - t0
    return {"grid": [n, n], "steps": steps, "precision": "f64",
            "gpu_s": round(tg, 4), "gpu_Gcell_updates_per_s": round(n * n * steps / tg / 1e9, 4),
            "cpu_oracle_all_cores_s": round(tp, 3), "cpu_threads": O.threads_parallel(),
            "cpu_oracle_1_thread_Gcell_updates_per_s": round(n * n * cpu_steps / t1 / 1e9, 5),
            "max_abs_diff_phi": float(np.max(np.abs(gphi - ref.phi))),
            "max_abs_diff_T": float(np.max(np.abs(gT - ref.T))),
            "solid_fraction": float((gphi > 0.5).mean())}


def dense_run(kb, size: int, nsteps: int):
    """The same trajectory with KG_KERNEL_FUSED_DENSE: the fused launch with
    no tile status (every tile loads and stores all of its fields; bitwise
    the same results, tests/test_gpu_tile_status.py).  Shows how much of the
    headline comes from skipping work the status proves unnecessary."""
    with kb.Solver(size, size, "f32", kb.default_params(**WORKLOAD)) as s:
        s.set_kernel("fused_dense")
        s.reset_seeds(**SEEDS)
        s.synchronize()
        ms = s.step_timed(nsteps)
    per = ms / nsteps / 1e3
    ach = ALG_BYTES_FP32 * size * size / per / 1e9
    return {"value": round(size * size / per / 1e9, 3), "unit": "Gcell-updates/s",
            "avg_launch_us": round(per * 1e6, 3), "achieved_GBs": round(ach, 1),
            "frac": round(ach / HBM_PEAK_GBS, 4)}


def fp64_error(kb, size: int, nsteps: int = 100):
    """Relative L2 error of the FP32 trajectory against the FP64 kernel after
    100 steps (the FP64 kernel matches the FP64 oracle to 1e-10 in the tests)."""
    out = {}
    fields = {}
    for prec in ("f32", "f64"):
        with kb.Solver(size, size, prec, kb.default_params(**WORKLOAD)) as s:
            s.reset_seeds(**SEEDS)
            s.step(nsteps)
            fields[prec] = (s.phase().astype(np.float64), s.temperature().astype(np.float64))
    for k, name in enumerate(("phi", "T")):
        a, b = fields["f32"][k], fields["f64"][k]
        out[name] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    return out


def calibrate_copy():
    """Copy 2 x 64 MiB (the FP32 phi and T of a 4096^2 grid) 20 times."""
    import torch
    n = 2 * 4096 * 4096
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(20):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read_per_copy": 4 * n,
                      "bytes_written_per_copy": 4 * n, "copies": 20}), flush=True)


def batched_vs_looped(kb, nsteps: int, looped: bool = True):
    """The ensemble from its initial state (disk r = 5 at every member's
    centre) for nsteps steps: one batched launch per step vs each member's
    own plain handle stepped in turn (every handle has its own stream, so
    members overlap on the GPU).  Both timed from the first step to the
    last, states resident on the device."""
    from crystalgrowth_amd import configs
    sizes, params = configs.c2_members()
    cells = sum(a * b for a, b in sizes)
    with kb.Batch(sizes, params, "f32") as b:
        b.reset_disk(-1, -1, -1, 5.0)
        mode, g = b.autotune(32)          # restores the state it ran from
        b.reset_disk(-1, -1, -1, 5.0)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(nsteps)
        b.synchronize()
        tb = time.perf_counter() - t0
        sb = [b.stats(m)["solid_fraction"] for m in (0, len(sizes) // 2, len(sizes) - 1)]
    if not looped:
        return {"members": len(sizes), "cells": cells, "steps": nsteps,
                "batched_Gcell_updates_per_s": round(cells * nsteps / tb / 1e9, 3),
                "launch_mode": [mode, g], "batched_s": round(tb, 3)}
    solvers = [kb.Solver(nx, ny, "f32", p) for (nx, ny), p in zip(sizes, params)]
    for (nx, ny), s in zip(sizes, solvers):
        s.reset_disk(nx // 2, ny // 2, 5.0)
    for s in solvers:
        s.synchronize()
    t0 = time.perf_counter()
    for s in solvers:
        s.step(nsteps)
    for s in solvers:
        s.synchronize()
    tl = time.perf_counter() - t0
    sl = [solvers[m].stats()["solid_fraction"] for m in (0, len(sizes) // 2, len(sizes) - 1)]
    for s in solvers:
        s.close()
    return {"members": len(sizes), "cells": cells, "steps": nsteps,
            "batched_Gcell_updates_per_s": round(cells * nsteps / tb / 1e9, 3),
            "looped_Gcell_updates_per_s": round(cells * nsteps / tl / 1e9, 3),
            "speedup_vs_looped": round(tl / tb, 3), "launch_mode": [mode, g],
            "batched_s": round(tb, 3), "looped_s": round(tl, 3),
            "same_state": sb == sl,
            "note": "members are bitwise equal to plain handles (tests/test_gpu_batch.py); "
                    "same_state compares three members' solid fractions at the end"}


def adjoint_vs_forward(kb):
    from crystalgrowth_amd import configs
    c = configs.C3
    n = c["n"]
    with kb.Solver(n, n, "f32", configs.c3_params("target")) as s:
        s.reset_disk(n // 2, n // 2, 5.0)
        s.step(c["steps"])
        target = s.phase()
    with kb.Solver(n, n, "f32", configs.c3_params("guess")) as s:
        s.reset_disk(n // 2, n // 2, 5.0)
        s.adjoint(2 * c["every"], target, c["every"])  # allocate the work buffers
        s.reset_disk(n // 2, n // 2, 5.0)
        t0 = time.perf_counter()
        g, _, _ = s.adjoint(c["steps"], target, c["every"], want_T0=True)
        ta = time.perf_counter() - t0
        s.reset_disk(n // 2, n // 2, 5.0)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(c["steps"])
        s.synchronize()
        tf = time.perf_counter() - t0
    return {"grid": [n, n], "steps": c["steps"], "checkpoint_every": c["every"],
            "adjoint_s": round(ta, 4), "forward_s": round(tf, 4),
            "adjoint_over_forward": round(ta / tf, 2), "loss": g["loss"],
            "grads": {k: g[k] for k in ("d_delta", "d_K", "d_eps_bar", "d_tau")},
            "note": "gradients checked against central differences in tests/test_gpu_adjoint.py"}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this solver runs on one GPU: --gpus must be 1 (the grid "
                                   "is not sharded across GPUs)", "n_gpus": args.gpus}))
        sys.exit(2)

    import crystalgrowth_amd as kb
    if kb.device_count() < 1:
        raise SystemExit("bench: no GPU visible")

    n = args.size
    cells = n * n
    with kb.Solver(n, n, "f32", kb.default_params(**WORKLOAD)) as s:
        s.set_kernel(args.kernel)
        # warmup on the same workload, then restart the trajectory at step 0
        s.reset_seeds(**SEEDS)
        if args.warmup > 0:
            s.step(args.warmup)
        s.synchronize()
        s.reset_seeds(**SEEDS)
        torch_sync()
        s.synchronize()
        t0 = time.perf_counter()
        if args.sync_every > 0:
            ev_ms, done = 0.0, 0
            while done < args.steps:
                k = min(args.sync_every, args.steps - done)
                ev_ms += s.step_timed(k)
                done += k
        else:
            ev_ms = s.step_timed(args.steps)  # HIP events on the solver's stream
        s.synchronize()
        torch_sync()
        wall = time.perf_counter() - t0
        st = s.stats()

    ms_per_step = ev_ms / max(args.steps, 1)
    value = cells * args.steps / (wall) / 1e9
    launch_s = ms_per_step / 1e3
    achieved = ALG_BYTES_FP32 * cells / launch_s / 1e9 if launch_s > 0 else 0.0
    rec = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / max(args.steps, 1), 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"kobayashi-{n}x{n}-fp32-zero-flux-noise-9-seeds",
                   "nx": n, "ny": n, "kernel": args.kernel, **WORKLOAD,
                   "nuclei": SEEDS["count"], "nucleus_radius": SEEDS["r"],
                   "dx": 0.03, "dt": 1e-4, "tau": 3e-4, "eps_bar": 0.01,
                   "alpha": 0.9, "gamma": 10.0, "teq": 1.0},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(n),
                     "kernel": ("k_pass1+k_pass2" if args.kernel == "two_pass" else
                                "k_step_fused<float> (all chunks) or k_step_fused_listed<float> "
                                "(live chunks, heaviest first; every 7th launch also answers the liveness request)"),
                     "avg_launch_us": round(ms_per_step * 1e3, 3),
                     "alg_bytes_per_cell_update": ALG_BYTES_FP32},
        "final_state": {"solid_fraction": st["solid_fraction"], "sum_phi": st["sum_phi"],
                        "nonfinite": st["nonfinite"]},
    }
    if args.kernel == "fused" and not args.no_dense:
        rec["dense_kernel"] = dense_run(kb, n, args.steps)
    if not args.no_fp64_check:
        rec["fp64_rel_err_100_steps"] = fp64_error(kb, n)
    if args.batch_steps > 0:
        rec["batched"] = batched_vs_looped(kb, args.batch_steps, not args.no_looped)
    if not args.no_adjoint:
        rec["adjoint"] = adjoint_vs_forward(kb)
    rec["cpu_baseline"] = cpu_baseline(n, args.cpu_steps) if args.cpu_steps > 0 else None
    if args.cpu_steps > 0:
        rec["cpu_all_cores"] = cpu_all_cores(n, 6 * args.cpu_steps)
        rec["reference_config"] = reference_config(kb, 20 * args.cpu_steps)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
