"""Benchmark of the Kobayashi step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--start S] [--size N]

Workload (the single-large-grid configuration): 4096 x 4096 FP32 grid,
zero-flux boundaries, dx = 0.03, dt = 1e-4, tau = 3e-4, eps_bar = 0.01,
delta = 0.04, j = 6, theta0 = 0.2, alpha = 0.9, gamma = 10, K = 2.0, Teq = 1,
noise a = 0.01, nine r = 5 nuclei at Philox-drawn centres (seed 0, margin 8),
a 20000-step horizon.

The headline step is one explicit-Euler step of the whole grid in which every
cell is loaded, updated and stored: the fused kernel without tile status
(k_step_dense).  The state is first advanced, untimed, to step S of the
trajectory (default 10000, mid-horizon: growing dendrites, the front, and the
quiescent far field together), then W untimed and K timed dense steps follow,
with the fields resident in HBM.  `value` is cell-updates per second over the
K timed steps (wall clock between two synchronisations); `roofline` prices
the average duration of the dense launch (HIP events on the solver's own
stream, one launch per step) at 16 algorithmic bytes per cell-update
(phi and T read once and written once) against the 8 TB/s HBM peak.
`cpu_baseline` times the CPU oracle (the reference's arithmetic restated in
C, one thread) on a bounded sample of the same workload.

Beside the headline, labelled apart:
  trajectory   -- the whole 20000-step horizon from the initial state, with
                  the default kernel (tile status: exact skips of work the
                  status proves unneeded, bitwise the dense results) and
                  with the dense kernel
  batched      -- the 512-member ensemble (sides 128..1024, per-member
                  parameters) over its 5000-step horizon from disk nuclei: one
                  batched launch per step (autotuned launch form and chunk shape) vs looping the
                  plain solver per member (one handle and stream per member)
  adjoint      -- loss + gradients of 1000 steps on 512^2 (checkpoint every 50)
                  vs the plain forward of the same steps
  fit          -- the 64-member batched parameter fit (500 steps, 50
                  batched-BFGS iterations): recovered-parameter errors, and one
                  batched gradient against the same gradient looped over plain
                  handles
  ceilings     -- what bounds the headline on this GPU: a device copy of the
                  dense step's algorithmic bytes (read 2 x 64 MiB, write
                  2 x 64 MiB; four buffer pairs in turn, so that sources are
                  not served from the last-level cache) and the dense launch
                  on the same grid with every tile flat (all fields +0), both
                  priced like the headline
  reference_config -- the reference's configuration (300^2 FP64, 2000 steps)
                  on the GPU and on the CPU oracle, compared cell by cell
  cpu_all_cores -- the oracle on OpenMP threads
--headline-only skips everything but the headline (for profiler runs);
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
    ap.add_argument("--steps", type=int, default=200, help="timed dense steps (K)")
    ap.add_argument("--warmup", type=int, default=20, help="untimed dense steps before them (W)")
    ap.add_argument("--start", type=int, default=10000,
                    help="trajectory step the dense window starts from (S; reached untimed)")
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="steps of the CPU oracle sample (0 skips the CPU baseline)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline window (for profiler runs)")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--horizon", type=int, default=20000,
                    help="steps of the trajectory legs (0 skips them)")
    ap.add_argument("--batch-steps", type=int, default=5000,
                    help="steps of the batched-vs-looped comparison (0 skips it)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--fit-iterations", type=int, default=50,
                    help="Adam iterations of the batched fit (0 skips it)")
    ap.add_argument("--no-looped", action="store_true",
                    help="batched leg without the looped plain handles (for profiling)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched leg in launch mode with 2 x 2 chunks, the tuner's pick on this "
                         "ensemble, without running the tuner (for profiling: the tuner's graph "
                         "candidates add captured launches to the trace)")
    ap.add_argument("--calibrate-copy", action="store_true")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def traffic_from_profiles(size: int):
    """HBM bytes per launch of the dense step measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE of k_step_dense over the headline window, scaled
    by the copy-kernel calibration) and recorded under profiles/; None when no
    measurement exists for this size."""
    path = os.path.join(ROOT, "profiles", f"k_step_dense_{size}x{size}_fp32_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(state, nsteps: int, start: int):
    """The CPU oracle (one thread) for nsteps steps from the state the GPU
    reached at trajectory step `start` (the state right after the headline
    window: reading the fields back before the timed steps slowed them)."""
    from oracle import oracle as O
    p = O.default_params(**WORKLOAD)
    size = state.phi.shape[0]
    t0 = time.perf_counter()
    O.run(state, p, nsteps, step0=start)
    dt = time.perf_counter() - t0
    cups = state.phi.size * nsteps / dt
    return {"value": cups / 1e9, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle (reference float arithmetic restated in C, gcc -O2, 1 thread) "
                      f"on the {size}x{size} workload, steps {start}-{start + nsteps - 1} from the "
                      f"GPU's state at step {start}, {dt:.1f} s on '{cpu_model()}' "
                      f"({os.cpu_count()} cores)"}


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


def trajectory(kb, size: int, nsteps: int):
    """The whole horizon from the initial state, timed with HIP events: the
    default kernel (tile status and launches of the live chunks only, exact:
    bitwise the dense results, tests/test_gpu_tile_status.py) and the dense
    kernel.  Labelled apart from the headline: the default kernel skips
    loads, stores and whole chunks the status proves unneeded, so its rate is
    not priced against the roofline."""
    out = {"steps": nsteps}
    for name in ("fused", "fused_dense"):
        with kb.Solver(size, size, "f32", kb.default_params(**WORKLOAD)) as s:
            s.set_kernel(name)
            s.reset_seeds(**SEEDS)
            s.synchronize()
            ms = s.step_timed(nsteps)
            st = s.stats()
        per = ms / nsteps / 1e3
        key = "exact_skip" if name == "fused" else "dense"
        out[key] = {"Gcell_updates_per_s": round(size * size / per / 1e9, 3),
                    "avg_step_us": round(per * 1e6, 3),
                    "final_solid_fraction": st["solid_fraction"]}
    out["dense"]["alg_frac_of_hbm_peak"] = round(
        ALG_BYTES_FP32 * size * size / (out["dense"]["avg_step_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
    out["same_final_state"] = (out["exact_skip"]["final_solid_fraction"] ==
                               out["dense"]["final_solid_fraction"])
    return out


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


def ceilings(kb, size: int, launches: int = 50):
    """Two references for the headline's roofline fraction, measured on the
    same GPU: a device-to-device copy of phi and T (the dense step's
    algorithmic bytes: 2 fields read, 2 written) and the dense launch on an
    all-+0 grid, where every tile takes the flat path (one memory round trip,
    no pass 1).  Both priced at 16 B per cell like the headline."""
    import ctypes
    from crystalgrowth_amd._lib import lib
    hip = lib()  # the HIP runtime the solver itself uses
    vp = ctypes.c_void_p
    for name, args in (("hipMalloc", [ctypes.POINTER(vp), ctypes.c_size_t]), ("hipFree", [vp]),
                       ("hipMemcpyAsync", [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]),
                       ("hipEventCreate", [ctypes.POINTER(vp)]), ("hipEventDestroy", [vp]),
                       ("hipEventRecord", [vp, vp]), ("hipEventSynchronize", [vp]),
                       ("hipEventElapsedTime", [ctypes.POINTER(ctypes.c_float), vp, vp])):
        getattr(hip, name).argtypes = args
    cells = size * size
    alg = ALG_BYTES_FP32 * cells
    nbytes = 2 * 4 * cells  # phi and T, FP32
    # four source/destination pairs used in turn (1 GiB in all), so that a
    # copy's source is not still in the 256 MiB last-level cache from the
    # copy before it, as the dense step's inputs are not
    pairs = [(vp(), vp()) for _ in range(4)]
    e0, e1 = vp(), vp()

    def ok(rc):
        if rc != 0:
            raise RuntimeError(f"bench: HIP error {rc} in the copy ceiling")

    try:
        for src, dst in pairs:
            ok(hip.hipMalloc(ctypes.byref(src), nbytes))
            ok(hip.hipMalloc(ctypes.byref(dst), nbytes))
        ok(hip.hipEventCreate(ctypes.byref(e0)))
        ok(hip.hipEventCreate(ctypes.byref(e1)))
        for src, dst in pairs:
            ok(hip.hipMemcpyAsync(dst, src, nbytes, 3, None))  # hipMemcpyDeviceToDevice
        ok(hip.hipEventRecord(e0, None))
        for i in range(launches):
            src, dst = pairs[i % len(pairs)]
            ok(hip.hipMemcpyAsync(dst, src, nbytes, 3, None))
        ok(hip.hipEventRecord(e1, None))
        ok(hip.hipEventSynchronize(e1))
        ms = ctypes.c_float(0.0)
        ok(hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1))
        copy_us = ms.value * 1e3 / launches
    finally:
        for src, dst in pairs:
            for p in (src, dst):
                if p.value:
                    hip.hipFree(p)
        for e in (e0, e1):
            if e.value:
                hip.hipEventDestroy(e)
    with kb.Solver(size, size, "f32", kb.default_params(**WORKLOAD)) as s:
        s.set_kernel("fused_dense")
        z = np.zeros((size, size), np.float32)
        s.set_fields(z, z, z)
        s.step(5)
        flat_us = s.step_timed(launches) * 1e3 / launches
    out = {}
    for key, us in (("device_copy", copy_us), ("dense_all_flat", flat_us)):
        out[key] = {"us": round(us, 3), "alg_GBs": round(alg / (us * 1e-6) / 1e9, 1),
                    "frac_of_hbm_peak": round(alg / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    out["note"] = ("device_copy: hipMemcpyAsync device to device of 2 x 64 MiB, four buffer "
                   "pairs in turn (HIP events); "
                   "dense_all_flat: k_step_dense on a 4096^2 grid of +0 fields (HIP events on "
                   "the solver's stream)")
    return out


TUNE_AT = 1000


def batched_vs_looped(kb, nsteps: int, looped: bool = True, autotune: bool = True):
    """The ensemble from its initial state (disk r = 5 at every member's
    centre) for nsteps steps: one batched launch per step vs each member's
    own plain handle stepped in turn -- all queued at once (every handle has
    its own stream, so members overlap on the GPU), and one member at a
    time.  All timed from the first step to the last, states resident on
    the device."""
    from crystalgrowth_amd import configs
    sizes, params = configs.c2_members()
    cells = sum(a * b for a, b in sizes)
    with kb.Batch(sizes, params, "f32") as b:
        # tuned on the run's own state at step TUNE_AT, where quiescent, flat
        # and active tiles are all present (at step 0 nearly every tile is
        # quiescent, which favours a shape that the later steps do not)
        if autotune:
            b.reset_disk(-1, -1, -1, 5.0)
            b.step(TUNE_AT)
            mode, g = b.autotune(32)      # restores the state it ran from
        else:
            b.set_chunk_shape(2, 2)
            b.set_mode("launch")
            mode, g = b.mode()
        shape = b.chunk_shape
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
                "launch_mode": [mode, g], "chunk_shape": list(shape), "tuned_at_step": TUNE_AT if autotune else None,
                "batched_s": round(tb, 3)}
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
    # the plain solver launched per instance one instance at a time: each
    # member's whole run finishes before the next one starts
    for (nx, ny), s in zip(sizes, solvers):
        s.reset_disk(nx // 2, ny // 2, 5.0)
        s.synchronize()
    t0 = time.perf_counter()
    for s in solvers:
        s.step(nsteps)
        s.synchronize()
    ts = time.perf_counter() - t0
    for s in solvers:
        s.close()
    return {"members": len(sizes), "cells": cells, "steps": nsteps,
            "batched_Gcell_updates_per_s": round(cells * nsteps / tb / 1e9, 3),
            "looped_Gcell_updates_per_s": round(cells * nsteps / tl / 1e9, 3),
            "speedup_vs_looped": round(tl / tb, 3),
            "looped_sequential_Gcell_updates_per_s": round(cells * nsteps / ts / 1e9, 3),
            "speedup_vs_looped_sequential": round(ts / tb, 3),
            "launch_mode": [mode, g], "chunk_shape": list(shape), "tuned_at_step": TUNE_AT if autotune else None,
            "batched_s": round(tb, 3), "looped_s": round(tl, 3), "looped_sequential_s": round(ts, 3),
            "same_state": sb == sl,
            "note": "looped: every member's own handle and stream, all members queued at once "
                    "(up to four overlap on the GPU's hardware queues); looped_sequential: the "
                    "same handles, each run to the horizon and synchronised before the next "
                    "starts; members are bitwise equal to plain handles "
                    "(tests/test_gpu_configs.py); same_state compares three members' solid "
                    "fractions at the end"}


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


def fit_summary(kb, iterations: int):
    """The 64-member batched parameter fit and one gradient batched vs looped."""
    from crystalgrowth_amd import configs, fit
    c = configs.C4
    sizes, hidden, guess = configs.c4_members(c["count"])
    t0 = time.perf_counter()
    res = fit.fit(sizes, hidden, guess, c["steps"], iterations, c["every"])
    tf = time.perf_counter() - t0
    targets = fit.make_targets(sizes, hidden, c["steps"], "f32")
    tb = fit.batched_adjoint_seconds(sizes, guess, targets, c["steps"], c["every"])
    tl = fit.looped_adjoint_seconds(sizes, guess, targets, c["steps"], c["every"])
    return {"members": len(sizes), "steps": c["steps"], "iterations": iterations,
            "fit_s": round(tf, 2), "loss_first": res["loss_first"], "loss_last": res["loss_last"],
            "param_error_start": res["param_error_start"],
            "param_error_end": res["param_error_end"],
            "one_gradient_batched_s": round(tb, 4), "one_gradient_looped_s": round(tl, 4),
            "batched_speedup_vs_looped": round(tl / tb, 3)}


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
    first = args.start + args.warmup
    # the headline solver stays open until the CPU sample is read from it at
    # the end: a readback of its fields slowed the dense launches that
    # followed by 3-15 % (3 x 64 MiB device-to-host, pageable or pinned
    # alike; without one they hold 52.2 us per launch), so no leg runs
    # after one
    s = kb.Solver(n, n, "f32", kb.default_params(**WORKLOAD))
    # reach step S of the trajectory untimed (default kernel: bitwise the
    # dense results), then W untimed and K timed dense steps
    s.reset_seeds(**SEEDS)
    s.step(args.start)
    s.synchronize()
    s.set_kernel("fused_dense")
    if args.warmup > 0:
        s.step(args.warmup)
    # the solver's stream is the only GPU work in this process: its
    # synchronisation brackets the timed steps (torch is not loaded here;
    # loading it beside the solver's runtime lengthened each launch by
    # about 5 us)
    s.synchronize()
    t0 = time.perf_counter()
    ev_ms = s.step_timed(args.steps)  # HIP events on the solver's stream
    s.synchronize()
    wall = time.perf_counter() - t0
    st = s.stats()

    per_launch = ev_ms / 1e3 / max(args.steps, 1)
    value = cells * args.steps / wall / 1e9
    achieved = ALG_BYTES_FP32 * cells / per_launch / 1e9 if per_launch > 0 else 0.0
    frac = achieved / HBM_PEAK_GBS
    if frac > 1.0:
        print(f"bench: roofline fraction {frac:.3f} above 1 -- the timed launches did not move "
              f"the algorithmic bytes", file=sys.stderr, flush=True)
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
        "config": {"workload": f"kobayashi-{n}x{n}-fp32-zero-flux-noise-9-seeds-dense-steps-"
                               f"{first}-{first + args.steps - 1}",
                   "nx": n, "ny": n, "kernel": "k_step_dense (every cell loaded, updated, stored)",
                   "window_start_step": first, "horizon": 20000, **WORKLOAD,
                   "nuclei": SEEDS["count"], "nucleus_radius": SEEDS["r"],
                   "dx": 0.03, "dt": 1e-4, "tau": 3e-4, "eps_bar": 0.01,
                   "alpha": 0.9, "gamma": 10.0, "teq": 1.0},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(frac, 4),
                     "traffic": traffic_from_profiles(n),
                     "kernel": "k_step_dense<float, zero-flux>: one launch per step, every tile",
                     "avg_launch_us": round(per_launch * 1e6, 3),
                     "alg_bytes_per_cell_update": ALG_BYTES_FP32,
                     "alg_bytes_per_launch": ALG_BYTES_FP32 * cells,
                     "frac_at_most_1": frac <= 1.0},
        "final_state": {"step": first + args.steps, "solid_fraction": st["solid_fraction"],
                        "sum_phi": st["sum_phi"], "nonfinite": st["nonfinite"]},
    }
    if args.headline_only:
        s.close()
        print(json.dumps(rec), flush=True)
        return
    rec["ceilings"] = ceilings(kb, n)
    if args.horizon > 0:
        rec["trajectory"] = trajectory(kb, n, args.horizon)
    if args.batch_steps > 0:
        rec["batched"] = batched_vs_looped(kb, args.batch_steps, not args.no_looped,
                                           not args.no_autotune)
    if not args.no_adjoint:
        rec["adjoint"] = adjoint_vs_forward(kb)
    if args.fit_iterations > 0:
        rec["fit"] = fit_summary(kb, args.fit_iterations)
    if not args.no_fp64_check:  # reads 4096^2 fields back: after the timed legs
        rec["fp64_rel_err_100_steps"] = fp64_error(kb, n)
    start_state = None
    if args.cpu_steps > 0:
        from oracle import oracle as O
        start_state = O.State(*s.fields())  # the state right after the timed steps
    s.close()
    rec["cpu_baseline"] = (cpu_baseline(start_state, args.cpu_steps, first + args.steps)
                           if start_state is not None else None)
    if args.cpu_steps > 0:
        rec["cpu_all_cores"] = cpu_all_cores(n, 6 * args.cpu_steps)
        rec["reference_config"] = reference_config(kb, 20 * args.cpu_steps)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
