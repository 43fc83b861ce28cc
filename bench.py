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
region; `roofline` prices the fused step kernel's average launch (HIP events
on the solver's own stream) at 16 algorithmic bytes per cell-update against
the 8 TB/s HBM peak; `cpu_baseline` times the CPU oracle (the reference's
arithmetic restated in C, one thread) on a bounded sample of the same
workload.
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
                seed=0)
SEEDS = dict(count=9, r=5.0, margin=8, seed=0)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--kernel", default="fused", choices=["fused", "two_pass"])
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="steps of the CPU oracle sample (0 skips the CPU baseline)")
    ap.add_argument("--no-fp64-check", action="store_true")
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


def main():
    args = parse()
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
                     "kernel": "k_step_fused<float>" if args.kernel == "fused" else "k_pass1+k_pass2",
                     "avg_launch_us": round(ms_per_step * 1e3, 3),
                     "alg_bytes_per_cell_update": ALG_BYTES_FP32},
        "final_state": {"solid_fraction": st["solid_fraction"], "sum_phi": st["sum_phi"],
                        "nonfinite": st["nonfinite"]},
    }
    if not args.no_fp64_check:
        rec["fp64_rel_err_100_steps"] = fp64_error(kb, n)
    rec["cpu_baseline"] = cpu_baseline(n, args.cpu_steps) if args.cpu_steps > 0 else None
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
