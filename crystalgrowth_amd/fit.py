"""Batched parameter fit of Kobayashi runs to target morphologies.

Each member m of an ensemble has hidden parameters; its target is phi after
`steps` steps from a disk nucleus.  Starting from a common guess, every
iteration resets all members to the same initial state (same noise counters),
runs the batched forward + adjoint (kg_batch_adjoint) for
L_m = sum (phi_m - target_m)^2, and takes one Adam step per member on the
fitted parameters.  `looped_adjoint_seconds` times the same gradient
evaluation done member by member with plain handles (kg_adjoint), the
baseline the batched path is compared with.

    python -m crystalgrowth_amd.fit [--members 64] [--iterations 50]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import configs
from .kobayashi import Batch, Solver

NAMES = ("delta", "K", "theta0")
LR = {"delta": 1e-3, "K": 0.03, "theta0": 0.03}
BOUNDS = {"delta": (0.0, 0.2), "K": (0.2, 4.0), "theta0": (-np.pi, np.pi)}


def lr_at(lr: dict, it: int, iterations: int, schedule: str) -> dict:
    """Learning rates of iteration `it` (1-based): constant, or a cosine decay
    to a tenth of the initial rate over the run."""
    if schedule == "constant" or iterations <= 1:
        return dict(lr)
    f = 0.1 + 0.9 * 0.5 * (1.0 + np.cos(np.pi * (it - 1) / (iterations - 1)))
    return {k: v * f for k, v in lr.items()}


def make_targets(sizes, hidden, steps, precision, r=5.0):
    with Batch(sizes, hidden, precision) as b:
        b.reset_disk(-1, -1, -1, r)
        b.step(steps)
        return [b.phase(m) for m in range(len(sizes))]


def fit(sizes, hidden, guess, steps, iterations, every=50, precision="f32", r=5.0, log=None,
        lr=None, schedule="constant"):
    """Adam on (delta, K, theta0) of every member; `lr` maps each name to its
    initial learning rate (default LR), `schedule` is "constant" or "cosine"."""
    lr = dict(LR if lr is None else lr)
    targets = make_targets(sizes, hidden, steps, precision, r)
    cur = [g.replace() for g in guess]
    M = len(sizes)
    m1 = np.zeros((M, len(NAMES)))
    m2 = np.zeros((M, len(NAMES)))
    b1, b2, eps = 0.9, 0.999, 1e-12
    losses = []
    t_adj = 0.0
    with Batch(sizes, cur, precision) as b:
        for it in range(1, iterations + 1):
            for m in range(M):
                b.set_params(m, cur[m])
            b.reset_disk(-1, -1, -1, r)
            t0 = time.perf_counter()
            grads, _ = b.adjoint(steps, targets, every)
            t_adj += time.perf_counter() - t0
            losses.append(float(sum(g["loss"] for g in grads)))
            rate = lr_at(lr, it, iterations, schedule)
            for m in range(M):
                g = np.array([grads[m]["d_" + n] for n in NAMES])
                m1[m] = b1 * m1[m] + (1 - b1) * g
                m2[m] = b2 * m2[m] + (1 - b2) * g * g
                mh = m1[m] / (1 - b1 ** it)
                vh = m2[m] / (1 - b2 ** it)
                upd = {}
                for k, n in enumerate(NAMES):
                    v = getattr(cur[m], n) - rate[n] * mh[k] / (np.sqrt(vh[k]) + eps)
                    lo, hi = BOUNDS[n]
                    upd[n] = float(min(max(v, lo), hi))
                cur[m] = cur[m].replace(**upd)
            if log:
                log(it, losses[-1])
    return {"members": M, "steps": steps, "iterations": iterations, "loss_first": losses[0],
            "loss_last": losses[-1], "param_error_start": param_errors(guess, hidden),
            "param_error_end": param_errors(cur, hidden), "batched_adjoint_seconds": t_adj,
            "fitted": [{n: getattr(c, n) for n in NAMES} for c in cur]}


def param_errors(params, hidden):
    """Mean relative error of delta and K, and mean absolute error (radians) of
    theta0 modulo 2 pi / j -- the period of eps(theta - theta0)."""
    out = {}
    for n in ("delta", "K"):
        out[n + "_rel"] = float(np.mean([abs(getattr(p, n) - getattr(h, n)) / abs(getattr(h, n))
                                         for p, h in zip(params, hidden)]))
    d = []
    for p, h in zip(params, hidden):
        per = 2 * np.pi / h.aniso
        e = (p.theta0 - h.theta0) % per
        d.append(min(e, per - e))
    out["theta0_abs_rad"] = float(np.mean(d))
    return out


def looped_adjoint_seconds(sizes, params, targets, steps, every=50, precision="f32", r=5.0):
    """One gradient evaluation of every member with plain handles, one after another."""
    total = 0.0
    for (nx, ny), p, t in zip(sizes, params, targets):
        with Solver(nx, ny, precision, p) as s:
            s.reset_disk(nx // 2, ny // 2, r)
            s.adjoint(min(2 * every, steps), t, every)  # allocate outside the timing
            s.reset_disk(nx // 2, ny // 2, r)
            t0 = time.perf_counter()
            s.adjoint(steps, t, every)
            total += time.perf_counter() - t0
    return total


def batched_adjoint_seconds(sizes, params, targets, steps, every=50, precision="f32", r=5.0):
    with Batch(sizes, params, precision) as b:
        b.reset_disk(-1, -1, -1, r)
        b.adjoint(min(2 * every, steps), targets, every)
        b.reset_disk(-1, -1, -1, r)
        t0 = time.perf_counter()
        b.adjoint(steps, targets, every)
        return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=configs.C4["count"])
    ap.add_argument("--iterations", type=int, default=configs.C4["iterations"])
    ap.add_argument("--steps", type=int, default=configs.C4["steps"])
    args = ap.parse_args()
    sizes, hidden, guess = configs.c4_members(args.members)
    res = fit(sizes, hidden, guess, args.steps, args.iterations, configs.C4["every"])
    targets = make_targets(sizes, hidden, args.steps, "f32")
    res["one_gradient_batched_s"] = batched_adjoint_seconds(sizes, guess, targets, args.steps)
    res["one_gradient_looped_s"] = looped_adjoint_seconds(sizes, guess, targets, args.steps)
    res["batched_speedup_vs_looped"] = res["one_gradient_looped_s"] / res["one_gradient_batched_s"]
    res.pop("fitted")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
